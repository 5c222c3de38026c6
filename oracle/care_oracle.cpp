// care_oracle: CPU restatement of CARE's read-correction path (care-cpu).
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle that the GPU
// product is checked against and the "port" CPU baseline of the benchmark.
// Nothing in the product (care_amd/, include/) may link, load or call it.
//
// It restates, in plain C++17 written for this project, the algorithm of the
// reference CPU corrector.  Every function names the reference file:line it
// follows.  It is pinned by closed forms, known answers and brute force in
// tests/test_oracle.py and tests/test_oracle_pins.py (codecs, quality bins,
// murmur finalizer against the published MurmurHash3 verification value,
// brute-force minhash and tables, character-level shifted Hamming,
// hand-built pileups for the MSA, refinement, HQ, filter, edit, forest and
// combine rules).  The reference care-cpu is not buildable here: its
// include/helpers points into an un-vendored git submodule.
//
// Build (see oracle/build.sh): g++ -O2 -std=c++17 -fopenmp -ffp-contract=off

#include "care_oracle.h"

#include <sys/stat.h>
#include <zlib.h>
#include <omp.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

using u32 = std::uint32_t;
using u64 = std::uint64_t;

// ---------------------------------------------------------------------------
// Options (defaults: include/options.hpp:36-98)
// ---------------------------------------------------------------------------
struct Options {
    int min_overlap = 30;
    float maxErrorRate = 0.2f;
    float min_overlap_ratio = 0.30f;
    bool excludeAmbiguousReads = false;
    bool correctCandidates = false;
    bool useQualityScores = false;
    bool outputCorrectionQualityLabels = false;
    float estimatedCoverage = 1.0f;
    float estimatedErrorrate = 0.06f;
    float m_coverage = 0.6f;
    int new_columns_to_correct = 15;
    int kmerlength = 20;
    int numHashFunctions = 48;
    int correctionType = 0;      // 0 classic, 1 forest, 3 learned gate (engine extension)
    int candidateLimit = 0;      // top candidates by map hits (0 = all, the reference)
    int maxRefinements = 5;      // different-region refinement rounds (the reference: 5)
    bool gateRefine = false;     // correctionType 3 with refinement: the plain method's decision columns
    int gateConstant = 1;        // correctionType 3: the oracle's gate accepts (1) or rejects (-1) every column
    std::string gateDump;        // correctionType 3: binary record of every decision column
    int gateConstantCands = 1;   // correctionTypeCands 3: the same for candidate columns
    std::string gateDumpCands;   // correctionTypeCands 3: binary record of every candidate decision column
    int correctionTypeCands = 0;
    float thresholdAnchor = 0.5f;
    float thresholdCands = 0.5f;
    float pairedFilterThreshold = 0.06f;
    int threads = 1;
    std::size_t fixedNumberOfReads = 0;
    int qualityScoreBits = 8;
    bool pairedEnd = false;
    std::string outputdirectory;
    std::vector<std::string> inputfiles;
    std::vector<std::string> outputfilenames;
    std::string forestAnchor, forestCands;
    std::uint32_t maxTreesAnchor = 0xFFFFFFFFu, maxTreesCands = 0xFFFFFFFFu;
    std::string dumpFile;
};

// ---------------------------------------------------------------------------
// Float comparison helpers (include/hostdevicefunctions.cuh:9-29)
// ---------------------------------------------------------------------------
inline bool feq(float l, float r) {
    const float d = l - r < 0 ? -(l - r) : l - r;
    return d < 1e-5f;
}
inline bool fleq(float l, float r) { return (l < r) || feq(l, r); }
inline bool fgeq(float l, float r) { return (l > r) || feq(l, r); }

// Overlap weight (include/hostdevicefunctions.cuh:93-97)
inline float overlapWeight(int nOps, int overlap, float maxMismatchRatio) {
    return 1.0f - std::sqrt(float(nOps) / (float(overlap) * maxMismatchRatio));
}

// Quality weight table (include/qualityscoreweights.hpp:15-30)
struct QualityWeights {
    float w[256];
    QualityWeights() {
        for (int i = 0; i < 256; i++) {
            float p = 1.0f;
            if (i >= 33) p = std::pow(10.0f, float(-(i - 33)) / 10.0f);
            w[i] = std::max(0.001f, 1.0f - p);
        }
    }
    float operator()(char c) const { return w[(unsigned char)c]; }
};
const QualityWeights& qweights() {
    static QualityWeights q;
    return q;
}

// ---------------------------------------------------------------------------
// 2-bit codec (include/sequencehelpers.hpp:136-205, 267-316)
//   A=0 C=1 G=2 T=3, 16 bases per word, first base in bits 31..30,
//   last word left-aligned.
// ---------------------------------------------------------------------------
inline int numInts2Bit(int len) { return (len + 15) / 16; }
inline int numIntsHiLo(int len) { return 2 * ((len + 31) / 32); }

inline u32 encodeBase(char c) {
    switch (c) {
        case 'C': return 1;
        case 'G': return 2;
        case 'T': return 3;
        default: return 0;
    }
}
inline char decodeBase(u32 b) { return "ACGT"[b & 3]; }

void encode2Bit(u32* out, const char* s, int len) {
    const int n = numInts2Bit(len);
    for (int i = 0; i < n; i++) out[i] = 0;
    for (int i = 0; i < len; i++) out[i / 16] = (out[i / 16] << 2) | encodeBase(s[i]);
    const int left = 2 * (n * 16 - len);
    if (left > 0) out[n - 1] <<= left;
}
inline u32 baseAt(const u32* enc, int i) { return (enc[i / 16] >> (30 - 2 * (i % 16))) & 3u; }
void decode2Bit(char* out, const u32* enc, int len) {
    for (int i = 0; i < len; i++) out[i] = decodeBase(baseAt(enc, i));
}
void revcomp2Bit(u32* out, const u32* in, int len) {
    const int n = numInts2Bit(len);
    for (int i = 0; i < n; i++) out[i] = 0;
    for (int i = 0; i < len; i++) {
        const u32 b = 3u - baseAt(in, len - 1 - i);
        out[i / 16] |= b << (30 - 2 * (i % 16));
    }
}
char complementChar(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
        default: return c;
    }
}
void revcompChars(char* out, const char* in, int len) {
    for (int i = 0; i < len; i++) out[len - 1 - i] = complementChar(in[i]);
}

// HiLo planes (include/sequencehelpers.hpp:645-689): hi plane holds the high
// bit of every base, lo plane the low bit, 32 bases per word, first base in
// bit 31.
void toHiLo(u32* out, const u32* enc, int len) {
    const int words = (len + 31) / 32;
    u32* hi = out;
    u32* lo = out + words;
    for (int w = 0; w < words; w++) hi[w] = lo[w] = 0;
    for (int i = 0; i < len; i++) {
        const u32 b = baseAt(enc, i);
        hi[i / 32] |= ((b >> 1) & 1u) << (31 - (i % 32));
        lo[i / 32] |= (b & 1u) << (31 - (i % 32));
    }
}

// ---------------------------------------------------------------------------
// Quality compression (include/qualityscorecompression.hpp:28-146, 268-366)
// ---------------------------------------------------------------------------
inline int qualNumInts(int len, int bits) { return (bits * len + 31) / 32; }

void encodeQuality(u32* out, const char* q, int len, int bits) {
    if (bits == 8) {
        std::memcpy(out, q, len);
        return;
    }
    const int perInt = 32 / bits;
    const int n = qualNumInts(len, bits);
    static const int b1[2] = {13, 256};
    static const int b2[4] = {4, 13, 20, 256};
    for (int w = 0; w < n; w++) {
        u32 data = 0;
        const int pend = std::min(len, (w + 1) * perInt);
        for (int p = w * perInt; p < pend; p++) {
            const int Q = int(q[p]) - 33;
            const int nb = 1 << bits;
            for (int x = 0; x < nb; x++) {
                const int bound = bits == 1 ? b1[x] : b2[x];
                if (Q <= bound) {
                    data = (data << bits) | u32(x);
                    break;
                }
            }
        }
        if (w == n - 1) data <<= ((w + 1) * perInt - len) * bits;
        out[w] = data;
    }
}

void decodeQuality(char* out, const u32* enc, int len, int bits) {
    if (bits == 8) {
        std::memcpy(out, enc, len);
        return;
    }
    static const char q1[2] = {'(', '5'};
    static const char q2[4] = {'$', '(', '1', '9'};
    const int perInt = 32 / bits;
    for (int p = 0; p < len; p++) {
        const int w = p / perInt, pos = p % perInt;
        const u32 v = (enc[w] >> ((32 - bits) - pos * bits)) & ((1u << bits) - 1u);
        out[p] = bits == 1 ? q1[v] : q2[v];
    }
}

// ---------------------------------------------------------------------------
// Input parsing (include/kseqpp/kseqpp.hpp:56-118): header = whole line
// after '@'/'>', multi-line sequences, '+' line skipped.
// ---------------------------------------------------------------------------
struct Read {
    std::string header, seq, qual;
};

struct LineReader {
    gzFile f = nullptr;
    std::vector<char> buf;
    std::size_t beg = 0, end = 0;
    bool eof = false;
    explicit LineReader(const std::string& name) {
        f = gzopen(name.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + name);
        gzbuffer(f, 1 << 20);
        buf.resize(1 << 20);
    }
    ~LineReader() {
        if (f) gzclose(f);
    }
    int getc() {
        if (beg >= end) {
            if (eof) return -1;
            const int r = gzread(f, buf.data(), unsigned(buf.size()));
            if (r <= 0) {
                eof = true;
                return -1;
            }
            beg = 0;
            end = std::size_t(r);
        }
        return (unsigned char)buf[beg++];
    }
    // appends until '\n'; returns false at EOF with nothing read
    bool getline(std::string& s) {
        int c;
        bool any = false;
        while ((c = getc()) >= 0) {
            any = true;
            if (c == '\n') break;
            s.push_back(char(c));
        }
        if (!s.empty() && s.back() == '\r') s.pop_back();
        return any;
    }
};

struct FastxReader {
    LineReader r;
    int last = 0;
    explicit FastxReader(const std::string& n) : r(n) {}
    bool next(Read& rd) {
        int c = 0;
        if (last == 0) {
            while ((c = r.getc()) >= 0 && c != '>' && c != '@') {}
            if (c < 0) return false;
            last = c;
        }
        rd.header.clear();
        rd.seq.clear();
        rd.qual.clear();
        if (!r.getline(rd.header) && rd.header.empty()) return false;
        while ((c = r.getc()) >= 0 && c != '>' && c != '+' && c != '@') {
            if (c == '\n') continue;
            rd.seq.push_back(char(c));
            r.getline(rd.seq);
        }
        if (c == '>' || c == '@') last = c;
        if (c != '+') return true;  // FASTA
        while ((c = r.getc()) >= 0 && c != '\n') {}
        while (rd.qual.size() < rd.seq.size()) {
            if (!r.getline(rd.qual)) break;
        }
        last = 0;
        if (rd.qual.size() != rd.seq.size()) throw std::runtime_error("quality length mismatch");
        return true;
    }
};

bool isFastq(const std::string& name) {
    LineReader r(name);
    int c;
    while ((c = r.getc()) >= 0 && c != '>' && c != '@') {}
    return c == '@';
}

// Reads every input file into memory in read-id order.  Paired input given as
// two files is interleaved (include/readlibraryio.hpp:130-190).
std::vector<Read> readAll(const Options& opt) {
    std::vector<Read> reads;
    if (opt.pairedEnd && opt.inputfiles.size() == 2) {
        FastxReader a(opt.inputfiles[0]), b(opt.inputfiles[1]);
        Read r1, r2;
        while (a.next(r1) && b.next(r2)) {
            reads.push_back(r1);
            reads.push_back(r2);
        }
    } else {
        for (const auto& f : opt.inputfiles) {
            FastxReader a(f);
            Read r;
            while (a.next(r)) reads.push_back(r);
        }
    }
    return reads;
}

// ---------------------------------------------------------------------------
// Read store (include/chunkedreadstorageconstruction.hpp:67-92, 216-300)
// ---------------------------------------------------------------------------
struct ReadStore {
    std::size_t n = 0;
    int maxLen = 0, minLen = 0;
    int pitch = 0;    // u32 words per 2-bit row
    int qbits = 8;
    int qpitch = 0;   // u32 words per encoded quality row
    bool hasQual = false;
    bool pairedEnd = false;
    std::vector<u32> seqs;
    std::vector<int> lens;
    std::vector<std::uint8_t> ambig;
    std::vector<u32> quals;
    std::size_t numAmbig = 0;

    const u32* seq(std::size_t i) const { return seqs.data() + i * pitch; }
    const u32* qual(std::size_t i) const { return quals.data() + i * qpitch; }

    // gatherQualities (include/chunkedreadstorage.hpp:532-606): decodes
    // min(qpitch*32/bits, outPitch) characters of the stored row, including
    // the row's zero padding beyond the read length.
    void gatherQuality(char* out, std::size_t id, int outPitch) const {
        const int maxCompressed = qpitch * 32 / qbits;
        const int l = std::min(maxCompressed, outPitch);
        decodeQuality(out, qual(id), l, qbits);
    }
};

// Non-ACGT characters become A,C,G,T cyclically; the counter restarts with
// every parser batch of 65536 reads and with every input file.
ReadStore buildReadStore(std::vector<Read>& reads, const Options& opt, const std::vector<std::size_t>& fileStarts) {
    ReadStore rs;
    rs.n = reads.size();
    rs.pairedEnd = opt.pairedEnd;
    rs.hasQual = opt.useQualityScores;
    rs.qbits = opt.qualityScoreBits;
    int maxL = 0, minL = std::numeric_limits<int>::max();
    for (const auto& r : reads) {
        maxL = std::max(maxL, int(r.seq.size()));
        minL = std::min(minL, int(r.seq.size()));
    }
    if (reads.empty()) minL = 0;
    rs.maxLen = maxL;
    rs.minLen = minL;
    rs.pitch = numInts2Bit(maxL);
    rs.qpitch = qualNumInts(maxL, rs.qbits);
    rs.seqs.assign(rs.n * rs.pitch, 0);
    rs.lens.resize(rs.n);
    rs.ambig.assign(rs.n, 0);
    if (rs.hasQual) rs.quals.assign(rs.n * rs.qpitch, 0);

    constexpr std::size_t kBatch = 65536;
    std::vector<std::size_t> starts = fileStarts;
    starts.push_back(rs.n);
    for (std::size_t f = 0; f + 1 < starts.size(); f++) {
        for (std::size_t b = starts[f]; b < starts[f + 1]; b += kBatch) {
            int ncount = 0;
            const std::size_t e = std::min(starts[f + 1], b + kBatch);
            for (std::size_t i = b; i < e; i++) {
                std::string& s = reads[i].seq;
                bool amb = false;
                for (auto& c : s) {
                    if (c == 'a') c = 'A';
                    else if (c == 'c') c = 'C';
                    else if (c == 'g') c = 'G';
                    else if (c == 't') c = 'T';
                    else if (c != 'A' && c != 'C' && c != 'G' && c != 'T') {
                        c = "ACGT"[ncount];
                        ncount = (ncount + 1) % 4;
                        amb = true;
                    }
                }
                rs.ambig[i] = amb;
                rs.numAmbig += amb;
                rs.lens[i] = int(s.size());
                encode2Bit(rs.seqs.data() + i * rs.pitch, s.data(), int(s.size()));
                if (rs.hasQual) encodeQuality(rs.quals.data() + i * rs.qpitch, reads[i].qual.data(), int(s.size()), rs.qbits);
            }
        }
    }
    return rs;
}

// ---------------------------------------------------------------------------
// Minhash signatures (include/cpusequencehasher.hpp:88-125,
// include/sequencehelpers.hpp:844-926, hashers.cuh:129-137)
// ---------------------------------------------------------------------------
inline u64 murmur64(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// For every hash function i < h: min over canonical k-mers c of
// murmur64(c + i), masked to 2k bits.  Sequences shorter than k keep the
// all-ones start value (masked).
void signature(u64* out, const u32* enc, int len, int k, int h) {
    const u64 mask = ~u64(0) >> (2 * (32 - k));
    std::vector<u64> mins(h, ~u64(0));
    if (len >= k) {
        u64 fw = 0, rc = 0;
        for (int i = 0; i < len; i++) {
            const u64 b = baseAt(enc, i);
            fw = ((fw << 2) | b) & mask;
            rc = (rc >> 2) | ((3 - b) << (2 * (k - 1)));
            if (i >= k - 1) {
                const u64 canon = std::min(fw, rc);
                for (int f = 0; f < h; f++) mins[f] = std::min(mins[f], murmur64(canon + u64(f)));
            }
        }
    }
    for (int f = 0; f < h; f++) out[f] = mins[f] & mask;
}

// ---------------------------------------------------------------------------
// Hash tables: per map, key -> sorted read ids.  Keys holding fewer than 2
// or more than cap values keep no values (include/groupbykey.hpp:99-249,
// include/config.hpp:32-34, include/minhasherlimit.hpp:10-15).
// ---------------------------------------------------------------------------
int resultsPerMapThreshold(float coverage) {
    int r = int(coverage * 2.5f);
    r = std::min(r, 65535);
    return std::max(10, r);
}

struct Tables {
    int k = 20, h = 0, cap = 0;
    struct Map {
        std::unordered_map<u64, std::pair<u32, u32>> index;  // key -> (offset, count)
        std::vector<u32> values;
    };
    std::vector<Map> maps;

    void build(const ReadStore& rs, int k_, int h_, int cap_) {
        k = k_;
        h = h_;
        cap = cap_;
        maps.assign(h, Map{});
        std::vector<u64> sigs(rs.n * std::size_t(h));
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < rs.n; i++) signature(&sigs[i * h], rs.seq(i), rs.lens[i], k, h);
#pragma omp parallel for schedule(dynamic, 1)
        for (int m = 0; m < h; m++) {
            std::vector<std::pair<u64, u32>> kv(rs.n);
            for (std::size_t i = 0; i < rs.n; i++) kv[i] = {sigs[i * h + m], u32(i)};
            std::sort(kv.begin(), kv.end());
            Map& mp = maps[m];
            mp.index.reserve(rs.n / 2);
            std::size_t b = 0;
            while (b < kv.size()) {
                std::size_t e = b;
                while (e < kv.size() && kv[e].first == kv[b].first) e++;
                const std::size_t cnt = e - b;
                if (cnt >= 2 && cnt <= std::size_t(cap)) {
                    mp.index[kv[b].first] = {u32(mp.values.size()), u32(cnt)};
                    for (std::size_t j = b; j < e; j++) mp.values.push_back(kv[j].second);
                }
                b = e;
            }
        }
    }

    // ordinaryminhasher.hpp:135-221: values of all maps concatenated in map
    // order; reads shorter than k are not queried.
    void query(std::vector<u32>& out, const u32* enc, int len) const {
        out.clear();
        if (len < k) return;
        std::vector<u64> sig(h);
        signature(sig.data(), enc, len, k, h);
        for (int m = 0; m < h; m++) {
            auto it = maps[m].index.find(sig[m]);
            if (it == maps[m].index.end()) continue;
            const auto& v = maps[m].values;
            out.insert(out.end(), v.begin() + it->second.first, v.begin() + it->second.first + it->second.second);
        }
    }
};

// ---------------------------------------------------------------------------
// Shifted Hamming alignment (src/cpu_alignment.cpp:33-323)
//
// Shifts are tried in the order 0,1,..,La-minov then -1,..,-(Lc-minov).  For
// a shift with overlap ov the mismatch count over the overlap is m; it is
// accepted if m < min(int(ov*maxErrorRate), best - total + 2ov), i.e. its
// score m + total - 2ov beats the best so far.  Once that bound is <= 0 the
// direction stops (no later shift of that direction can be accepted, since
// the overlap never grows).  Result: first minimum in that order.
// ---------------------------------------------------------------------------
struct Alignment {
    int score = 0, overlap = 0, shift = 0, nOps = 0;
    bool valid = false;
};

// mismatches between a[sa .. sa+ov) and b[sb .. sb+ov) over HiLo planes
inline int hiloMismatch(const u32* ahi, const u32* alo, int sa, const u32* bhi, const u32* blo, int sb, int ov) {
    auto fetch = [](const u32* p, int bitpos, int nwords) -> u32 {
        const int w = bitpos >> 5, s = bitpos & 31;
        const u32 x = w < nwords ? p[w] : 0u;
        const u32 y = (w + 1) < nwords ? p[w + 1] : 0u;
        return s == 0 ? x : (x << s) | (y >> (32 - s));
    };
    (void)fetch;
    int m = 0;
    for (int off = 0; off < ov; off += 32) {
        const int n = std::min(32, ov - off);
        const u32 mask = n == 32 ? 0xFFFFFFFFu : (0xFFFFFFFFu << (32 - n));
        const int ia = sa + off, ib = sb + off;
        auto get = [](const u32* p, int bitpos) -> u32 {
            const int w = bitpos >> 5, s = bitpos & 31;
            return s == 0 ? p[w] : (p[w] << s) | (p[w + 1] >> (32 - s));
        };
        const u32 x = (get(ahi, ia) ^ get(bhi, ib)) | (get(alo, ia) ^ get(blo, ib));
        m += __builtin_popcount(x & mask);
    }
    return m;
}

Alignment align(const u32* anchorEnc, int la, const u32* candEnc, int lc, int minOverlap, float maxErrorRate,
                float minOverlapRatio) {
    // planes padded by one word so that get() may read p[w+1]
    const int wa = (la + 31) / 32, wc = (lc + 31) / 32;
    std::vector<u32> A(2 * (wa + 1), 0), C(2 * (wc + 1), 0);
    {
        std::vector<u32> t(2 * wa);
        toHiLo(t.data(), anchorEnc, la);
        std::copy(t.begin(), t.begin() + wa, A.begin());
        std::copy(t.begin() + wa, t.end(), A.begin() + wa + 1);
        std::vector<u32> u(2 * wc);
        toHiLo(u.data(), candEnc, lc);
        std::copy(u.begin(), u.begin() + wc, C.begin());
        std::copy(u.begin() + wc, u.end(), C.begin() + wc + 1);
    }
    const u32 *ahi = A.data(), *alo = A.data() + wa + 1, *chi = C.data(), *clo = C.data() + wc + 1;
    const int total = la + lc;
    const int minov = std::max(minOverlap, int(float(la) * minOverlapRatio));
    int best = total, bestShift = -lc;
    auto tryShift = [&](int shift, int ov) -> bool {
        const int lim = std::min(int(float(ov) * maxErrorRate), best - total + 2 * ov);
        if (lim <= 0) return false;
        const int m = shift >= 0 ? hiloMismatch(ahi, alo, shift, chi, clo, 0, ov)
                                 : hiloMismatch(ahi, alo, 0, chi, clo, -shift, ov);
        if (m < lim) {
            const int score = m + total - 2 * ov;
            if (score < best) {
                best = score;
                bestShift = shift;
            }
        }
        return true;
    };
    for (int s = 0; s < la - minov + 1; s++)
        if (!tryShift(s, std::min(la - s, lc))) break;
    for (int s = -1; s >= -lc + minov; s--)
        if (!tryShift(s, std::min(la, lc + s))) break;
    Alignment r;
    r.valid = bestShift != -lc;
    const int cb = std::max(-bestShift, 0);
    const int ce = std::min(lc, la - bestShift);
    r.overlap = ce - cb;
    r.shift = bestShift;
    r.score = best;
    r.nOps = best - total + 2 * r.overlap;
    return r;
}

// Orientation choice (include/alignmentorientation.hpp:14-92); maxErrorRate
// argument is 4 * estimatedErrorrate.  1 = forward, 2 = reverse complement,
// 3 = none.
int chooseOrientation(const Alignment& f, const Alignment& r, int la, float minOverlapRatio, int minOverlap,
                      float maxErr) {
    const int minov = int(la * minOverlapRatio) > minOverlap ? int(la * minOverlapRatio) : minOverlap;
    if (f.valid && f.overlap >= minov) {
        if (r.valid && r.overlap >= minov) {
            const float ratio = float(f.nOps) / f.overlap;
            const float rratio = float(r.nOps) / r.overlap;
            if (ratio < rratio) return ratio < maxErr ? 1 : 3;
            if (rratio < ratio) return rratio < maxErr ? 2 : 3;
            if (ratio < maxErr) return f.overlap > r.overlap ? 1 : 2;
            return 3;
        }
        return float(f.nOps) / f.overlap < maxErr ? 1 : 3;
    }
    if (r.valid && r.overlap >= minov && float(r.nOps) / r.overlap < maxErr) return 2;
    return 3;
}

// ---------------------------------------------------------------------------
// Forest classifier (include/forest.hpp:22-227): trees stored in preorder,
// node = att u8, split f64, flag u8 (bit1: left child is a leaf, bit0: right
// child is a leaf), leaves hold P(accept) as f32.  decide() averages leaf
// probabilities over trees and accepts as soon as the running mean reaches
// the threshold.
// ---------------------------------------------------------------------------
struct Forest {
    struct Node {
        std::uint8_t att = 0;
        std::uint8_t flag = 0;
        double split = 0;
        union {
            std::uint32_t idx;
            float prob;
        } lhs{}, rhs{};
    };
    std::vector<std::vector<Node>> trees;
    float thresh = 0.5f;

    template <class T>
    static T readT(std::istream& is) {
        T v;
        char buf[sizeof(T)];
        is.read(buf, sizeof(T));
        if (!is) throw std::runtime_error("truncated forest file");
        std::memcpy(&v, buf, sizeof(T));
        return v;
    }
    static void populate(std::istream& is, std::vector<Node>& tree) {
        const std::size_t me = tree.size();
        tree.emplace_back();
        tree[me].att = readT<std::uint8_t>(is);
        tree[me].split = readT<double>(is);
        tree[me].flag = readT<std::uint8_t>(is);
        if (tree[me].flag & 2) {
            tree[me].lhs.prob = readT<float>(is);
        } else {
            tree[me].lhs.idx = std::uint32_t(tree.size());
            populate(is, tree);
        }
        if (tree[me].flag & 1) {
            tree[me].rhs.prob = readT<float>(is);
        } else {
            tree[me].rhs.idx = std::uint32_t(tree.size());
            populate(is, tree);
        }
    }
    void load(const std::string& path, const std::string& expectedDesc, std::uint32_t maxTrees, float t) {
        std::ifstream is(path, std::ios::binary);
        if (!is) throw std::runtime_error("Loading classifier file failed! " + path);
        const auto len = readT<std::uint64_t>(is);
        std::string desc(len, '\0');
        is.read(&desc[0], std::streamsize(len));
        if (desc != expectedDesc)
            throw std::runtime_error("Classifier and extractor descriptors do not match! Expected: " + expectedDesc +
                                     " Received: " + desc);
        const auto n = readT<std::uint32_t>(is);
        trees.assign(std::min(n, maxTrees), {});
        for (auto& tree : trees) {
            tree.reserve(readT<std::uint32_t>(is));
            populate(is, tree);
        }
        thresh = t;
    }
    static float leaf(const std::vector<Node>& tree, const float* f) {
        std::size_t i = 0;
        for (;;) {
            const Node& n = tree[i];
            if (double(f[n.att]) <= n.split) {
                if (n.flag & 2) return n.lhs.prob;
                i = n.lhs.idx;
            } else {
                if (n.flag & 1) return n.rhs.prob;
                i = n.rhs.idx;
            }
        }
    }
    bool decide(const float* f) const {
        float sum = 0.f;
        for (const auto& tree : trees) {
            sum += leaf(tree, f);
            if (sum / float(trees.size()) >= thresh) return true;
        }
        return false;
    }
};

// ---------------------------------------------------------------------------
// MSA (src/msa.cpp:9-173)
// ---------------------------------------------------------------------------
struct MSA {
    int nColumns = 0, anchorBegin = 0, anchorEnd = 0, nCandidates = 0;
    std::vector<int> cA, cC, cG, cT, coverage, origCoverages;
    std::vector<float> wA, wC, wG, wT, support, origWeights;
    std::vector<char> consensus;

    struct Input {
        bool useQ;
        int anchorLength;
        int nCandidates;
        int candPitch, qualPitch;
        const char* anchor;
        const char* cands;
        const char* anchorQ;
        const char* candQ;
        const int* candLen;
        const int* shifts;
        const float* weights;
    } in{};

    void build(const Input& args) {
        in = args;
        nCandidates = args.nCandidates;
        int startindex = 0, endindex = args.anchorLength;
        for (int i = 0; i < nCandidates; i++) {
            startindex = std::min(args.shifts[i], startindex);
            endindex = std::max(args.candLen[i] + args.shifts[i], endindex);
        }
        nColumns = endindex - startindex;
        anchorBegin = std::max(-startindex, 0);
        anchorEnd = anchorBegin + args.anchorLength;
        for (auto* v : {&cA, &cC, &cG, &cT, &coverage, &origCoverages}) v->assign(nColumns, 0);
        for (auto* v : {&wA, &wC, &wG, &wT, &support, &origWeights}) v->assign(nColumns, 0.f);
        consensus.assign(nColumns, 0);
        const auto& qw = qweights();
        auto add = [&](const char* s, const char* q, int len, int shift, float f) {
            for (int i = 0; i < len; i++) {
                const int g = anchorBegin + shift + i;
                const float w = f * (args.useQ ? qw(q[i]) : 1.0f);
                switch (s[i]) {
                    case 'A': cA[g]++; wA[g] += w; break;
                    case 'C': cC[g]++; wC[g] += w; break;
                    case 'G': cG[g]++; wG[g] += w; break;
                    case 'T': cT[g]++; wT[g] += w; break;
                    default: break;
                }
                coverage[g]++;
            }
        };
        add(args.anchor, args.anchorQ, args.anchorLength, 0, 1.0f);
        for (int c = 0; c < nCandidates; c++)
            add(args.cands + std::size_t(c) * args.candPitch, args.useQ ? args.candQ + std::size_t(c) * args.qualPitch : nullptr,
                args.candLen[c], args.shifts[c], args.weights[c]);
        for (int col = 0; col < nColumns; col++) {
            char cons = 'A';
            float cw = wA[col];
            if (wC[col] > cw) { cons = 'C'; cw = wC[col]; }
            if (wG[col] > cw) { cons = 'G'; cw = wG[col]; }
            if (wT[col] > cw) { cons = 'T'; cw = wT[col]; }
            consensus[col] = cons;
            const float colw = wA[col] + wC[col] + wG[col] + wT[col];
            support[col] = cw / colw;
        }
        for (int col = anchorBegin; col < anchorEnd; col++) {
            switch (args.anchor[col - anchorBegin]) {
                case 'A': origWeights[col] = wA[col]; origCoverages[col] = cA[col]; break;
                case 'C': origWeights[col] = wC[col]; origCoverages[col] = cC[col]; break;
                case 'G': origWeights[col] = wG[col]; origCoverages[col] = cG[col]; break;
                case 'T': origWeights[col] = wT[col]; origCoverages[col] = cT[col]; break;
                default: break;
            }
        }
    }

    struct Props {
        float avg_support = 0, min_support = 0;
        int max_coverage = 0, min_coverage = 0;
        bool isHQ = false;
    };

    // src/msa.cpp:304-376
    Props properties(float err, float cov, float mcov) const { return properties(anchorBegin, anchorEnd, err, cov, mcov); }
    Props properties(int first, int last, float err, float cov, float mcov) const {
        Props p;
        const float avgT = 1.0f - 1.0f * err;
        // care-cpu (built -O3 -march=native) evaluates every 1 - 3x threshold
        // as one fused multiply-add
        const float minT = std::fma(-3.0f, err, 1.0f);
        const float covT = mcov / 6.0f * cov;
        const int dist = last - first;
        p.min_support = *std::min_element(support.begin() + first, support.begin() + last);
        float sum = 0.0f;
        for (int c = first; c < last; c++) sum += support[c];
        p.avg_support = sum / dist;
        p.min_coverage = *std::min_element(coverage.begin() + first, coverage.begin() + last);
        p.max_coverage = *std::max_element(coverage.begin() + first, coverage.begin() + last);
        const bool allGood = fgeq(p.avg_support, avgT) && fgeq(p.min_support, minT) && fgeq(float(p.min_coverage), covT);
        p.isHQ = false;
        if (allGood) {
            int smallest = 100;
            const int pct = int(std::ceil(err * 100.0f));
            for (int percent = pct; percent >= 0; percent--) {
                const float factor = percent / 100.0f;
                if (fgeq(p.avg_support, 1.0f - factor) && fgeq(p.min_support, std::fma(-3.0f, factor, 1.0f))) smallest = percent;
            }
            p.isHQ = fgeq(float(p.min_coverage), covT) && fleq(float(smallest), pct * 0.5f);
        }
        return p;
    }

    // src/msa.cpp:379-453
    std::string correctedAnchor(const Props& p, float err, float cov, float mcov) const {
        std::string out(in.anchorLength, 'A');
        const bool byConsensus = fgeq(p.avg_support, 1.0f - err) && fgeq(p.min_support, std::fma(-3.0f, err, 1.0f)) &&
                                 fgeq(float(p.min_coverage), mcov / 6.0f * cov);
        for (int i = 0; i < in.anchorLength; i++) {
            const int c = anchorBegin + i;
            if (byConsensus) out[i] = consensus[c];
            else out[i] = (support[c] > 0.90f && origCoverages[c] <= 2) ? consensus[c] : in.anchor[i];
        }
        return out;
    }

    // src/msa.cpp:532-734; returns true and per-candidate removal flags if a
    // column was found
    bool differentRegion(int datasetCoverage, std::vector<char>& remove) const {
        bool mismatch = false;
        for (int i = 0; i < in.anchorLength; i++)
            if (in.anchor[i] != consensus[anchorBegin + i]) { mismatch = true; break; }
        if (!mismatch) return false;
        const int sig = int(datasetCoverage * 0.3f);
        int col = -1, foundIdx = -1, consIdx = -1;
        for (int c = anchorBegin; c < anchorEnd && col < 0; c++) {
            const int counts[4] = {cA[c], cC[c], cG[c], cT[c]};
            const int ci = "\0\1\2\3"[0] + (consensus[c] == 'A' ? 0 : consensus[c] == 'C' ? 1 : consensus[c] == 'G' ? 2 : 3);
            int s = -1;
            for (int i = 0; i < 4; i++)
                if (i != ci && sig <= counts[i]) s = i;
            if (s != -1) {
                col = c;
                foundIdx = s;
                consIdx = ci;
            }
        }
        (void)consIdx;
        if (col < 0) return false;
        const char found = "ACGT"[foundIdx];
        const char orig = in.anchor[col - anchorBegin];
        const bool keepMatching = orig == found;
        remove.assign(nCandidates, 0);
        for (int c = 0; c < nCandidates; c++) {
            const int rb = anchorBegin + in.shifts[c];
            const int re = rb + in.candLen[c];
            if (col < rb || re <= col) continue;
            const char b = in.cands[std::size_t(c) * in.candPitch + (col - rb)];
            remove[c] = keepMatching ? (b != found) : (b == found);
        }
        for (int c = 0; c < nCandidates; c++) {
            if (remove[c] && in.weights[c] >= 0.90f) {
                std::fill(remove.begin(), remove.end(), 0);
                break;
            }
        }
        return true;
    }
};

// ---------------------------------------------------------------------------
// Single-end mismatch-ratio filter (include/corrector.hpp:1053-1149): the
// threshold is 2, 3 or 4 x the estimated error rate, the smallest whose
// count of candidates below it reaches coverage x m_coverage; with no
// candidate below 4x, or none reaching it at any level, nothing passes.
// ---------------------------------------------------------------------------
float seFilterThreshold(const float* ratios, int n, float estimatedErrorrate, float estimatedCoverage, float mCoverage) {
    const float base = estimatedErrorrate * 1.0f;
    const float goodT = estimatedCoverage * mCoverage;
    int counts[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        if (ratios[i] < 2 * base) counts[0]++;
        if (ratios[i] < 3 * base) counts[1]++;
        if (ratios[i] < 4 * base) counts[2]++;
    }
    float thr = std::numeric_limits<float>::lowest();
    if (counts[0] > 0 || counts[1] > 0 || counts[2] > 0) {
        if (counts[0] >= goodT) thr = 2 * base;
        else if (counts[1] >= goodT) thr = 3 * base;
        else if (counts[2] >= goodT) thr = 4 * base;
    }
    return thr;
}

// Candidates that candidate correction writes (MSA columns [qb, qe), anchor
// columns [ab, ae), ncc = new columns to correct): classic
// (src/msa.cpp:457-522) needs the start within +-ncc of the anchor start and
// the end at most ncc past the anchor end; the forest and the learned gate
// (include/corrector.hpp:1539-1623) only bound the start from below.
bool candidateInRange(int qb, int qe, int ab, int ae, int ncc, bool forestOrGate) {
    if (forestOrGate) return qb >= ab - ncc && qe <= ae + ncc;
    return ab - ncc <= qb && qb <= ab + ncc && qe <= ae + ncc;
}

// ---------------------------------------------------------------------------
// Per-read results and the combine rule (src/correctionresultoutput.cpp:730-1023)
// ---------------------------------------------------------------------------
struct Result {
    std::uint32_t readId = 0;
    bool isAnchor = false;
    bool hq = false;
    bool useEdits = false;
    std::vector<std::pair<int, char>> edits;
    std::string sequence;
};

// edits of corrected vs original, at most len/7 (include/corrector.hpp:1639-1761)
void makeEdits(Result& r, const std::string& corrected, const char* original, bool ambiguous) {
    r.sequence = corrected;
    if (ambiguous) {
        r.useEdits = false;
        return;
    }
    const int len = int(corrected.size());
    const int maxEdits = len / 7;
    int edits = 0;
    for (int i = 0; i < len && edits <= maxEdits; i++) {
        if (corrected[i] != original[i]) {
            r.edits.emplace_back(i, corrected[i]);
            edits++;
        }
    }
    r.useEdits = edits <= maxEdits;
}

// ---------------------------------------------------------------------------
// The corrector
// ---------------------------------------------------------------------------
struct AnchorDump {
    std::uint32_t id;
    std::vector<u32> candidates;        // after query/unique/self removal
    std::vector<int> flags;             // orientation per candidate
    std::vector<Alignment> chosen;      // alignment of the chosen orientation
    std::vector<u32> msaIds;            // candidates entering the final MSA
    std::string consensus;              // final MSA consensus over all columns
    int anchorBegin = 0;
    bool corrected = false, hq = false;
    std::string correctedSeq;
};

struct Corrector {
    const Options& opt;
    const ReadStore& rs;
    const Tables& tables;
    const Forest* forestA = nullptr;
    const Forest* forestC = nullptr;

    Corrector(const Options& o, const ReadStore& r, const Tables& t) : opt(o), rs(r), tables(t) {}

    std::vector<u32> candidatesOf(std::uint32_t id) const {
        std::vector<u32> ids;
        if (opt.excludeAmbiguousReads && rs.ambig[id]) return ids;
        tables.query(ids, rs.seq(id), rs.lens[id]);
        std::sort(ids.begin(), ids.end());
        if (opt.candidateLimit > 0) {
            // the engine's candidate limit (its gated mode and reduced
            // pipeline): keep the candidates hit by the most maps (ties to
            // the smaller id), after self / ambiguous removal
            std::vector<std::pair<int, u32>> hc;
            for (std::size_t i = 0; i < ids.size();) {
                std::size_t j = i;
                while (j < ids.size() && ids[j] == ids[i]) j++;
                if (ids[i] != id && !(opt.excludeAmbiguousReads && rs.ambig[ids[i]])) hc.push_back({int(j - i), ids[i]});
                i = j;
            }
            if (int(hc.size()) > opt.candidateLimit) {
                std::stable_sort(hc.begin(), hc.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
                hc.resize(std::size_t(opt.candidateLimit));
                std::sort(hc.begin(), hc.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
            }
            ids.clear();
            for (const auto& e : hc) ids.push_back(e.second);
            return ids;
        }
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        auto it = std::lower_bound(ids.begin(), ids.end(), id);
        if (it != ids.end() && *it == id) ids.erase(it);
        if (opt.excludeAmbiguousReads)
            ids.erase(std::remove_if(ids.begin(), ids.end(), [&](u32 x) { return rs.ambig[x] != 0; }), ids.end());
        return ids;
    }

    struct Cand {
        u32 id;
        int len;
        std::vector<u32> enc;  // oriented 2-bit
        Alignment al;
        int flag;
        bool paired;
        float weight;
        std::string chars;
        std::string qual;
    };

    struct GateColumn {
        u32 id;
        std::uint16_t pos;
        std::uint8_t orig, cons;
        float f[32];
    };

    struct Output {
        bool hasAnchor = false;
        Result anchor;
        std::vector<Result> cands;
        bool hq = false;
        std::vector<GateColumn> gate;
        std::vector<GateColumn> gateCands;
    };

    static int baseCode(char c) { return c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 0; }

    // The engine's 32 gate features of position i of a read in the MSA
    // starting at column `begin` (the anchor, or a candidate in its MSA
    // orientation with its orientation's qualities); p are the MSA
    // properties over the read's columns (no counterpart in the reference;
    // restated from care_amd/csrc/correct.hip gate_features).
    void gateFeatures(float* f, const MSA& msa, const MSA::Props& p, const std::string& anchor, const std::string& anchorQ,
                      int begin, int i, int nd, int m) const {
        const int col = begin + i;
        const int cols = msa.nColumns;
        const int la = int(anchor.size());
        const float estCov = opt.estimatedCoverage;
        const char o = anchor[i], c = msa.consensus[col];
        const float cov = float(msa.coverage[col]);
        const float wsum = msa.wA[col] + msa.wC[col] + msa.wG[col] + msa.wT[col];
        const float wo = weightOf(msa, o, col), wc = weightOf(msa, c, col);
        const float no = countOf(msa, o, col), nc = countOf(msa, c, col);
        auto mism = [&](int ii) { return msa.consensus[begin + ii] != anchor[ii]; };
        f[0] = no / cov;
        f[1] = wo / wsum;
        f[2] = nc / cov;
        f[3] = msa.support[col];
        f[4] = wo / std::max(no, 1.0f);
        f[5] = wc / std::max(nc, 1.0f);
        f[6] = cov / estCov;
        f[7] = wsum / estCov;
        f[8] = (cov - no - nc) / cov;
        f[9] = opt.useQualityScores ? qweights()(anchorQ[i]) : 1.0f;
        f[10] = float(i) / float(la);
        f[11] = float(std::min(i, la - 1 - i)) / float(la);
        f[12] = p.avg_support;
        f[13] = p.min_support;
        f[14] = float(p.min_coverage) / estCov;
        f[15] = float(p.max_coverage) / estCov;
        f[16] = float(m) / estCov;
        const int dcol[4] = {-2, -1, 1, 2};
        for (int k = 0; k < 4; k++) {
            const int cc = col + dcol[k];
            f[17 + k] = (cc >= 0 && cc < cols) ? msa.support[cc] : 0.0f;
            const int ii = i + dcol[k];
            f[21 + k] = (ii >= 0 && ii < la && mism(ii)) ? 1.0f : 0.0f;
        }
        f[25] = float(nd) / float(la);
        f[26] = col - 1 >= 0 ? float(msa.coverage[col - 1]) / estCov : 0.0f;
        f[27] = col + 1 < cols ? float(msa.coverage[col + 1]) / estCov : 0.0f;
        f[28] = (wc - wo) / wsum;
        const float q0 = float(msa.cA[col]) / cov, q1 = float(msa.cC[col]) / cov, q2 = float(msa.cG[col]) / cov,
                    q3 = float(msa.cT[col]) / cov;
        f[29] = q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3;
        f[30] = wsum / cov;
        int near = 0;
        for (int d = -5; d <= 5; d++) {
            const int ii = i + d;
            if (d != 0 && ii >= 0 && ii < la && mism(ii)) near++;
        }
        f[31] = float(near) / 10.0f;
    }

    // include/corrector.hpp:145-323 (process) and :325-536 (processMulti),
    // for one anchor whose candidate list and pair flags are given
    Output correctOne(std::uint32_t id, const std::vector<u32>& ids, const std::vector<char>& paired, AnchorDump* dump) const {
        Output out;
        const int la = rs.lens[id];
        const u32* aenc = rs.seq(id);
        std::string anchor(la, 'A');
        decode2Bit(&anchor[0], aenc, la);
        if (dump) {
            dump->id = id;
            dump->candidates = ids;
        }
        if (ids.empty()) return out;
        const int pitchQ = rs.maxLen;
        std::vector<Cand> cands;
        cands.reserve(ids.size());
        std::vector<u32> rc(rs.pitch);
        const float maxErr4 = opt.estimatedErrorrate * 4.0f;
        for (std::size_t i = 0; i < ids.size(); i++) {
            const u32 cid = ids[i];
            const int lc = rs.lens[cid];
            const u32* cenc = rs.seq(cid);
            revcomp2Bit(rc.data(), cenc, lc);
            const Alignment f = align(aenc, la, cenc, lc, opt.min_overlap, opt.maxErrorRate, opt.min_overlap_ratio);
            const Alignment r = align(aenc, la, rc.data(), lc, opt.min_overlap, opt.maxErrorRate, opt.min_overlap_ratio);
            const int flag = chooseOrientation(f, r, la, opt.min_overlap_ratio, opt.min_overlap, maxErr4);
            if (dump) {
                dump->flags.push_back(flag);
                dump->chosen.push_back(flag == 2 ? r : f);
            }
            if (flag == 3) continue;
            Cand c;
            c.id = cid;
            c.len = lc;
            c.flag = flag;
            c.al = flag == 1 ? f : r;
            c.enc.assign(flag == 1 ? cenc : rc.data(), (flag == 1 ? cenc : rc.data()) + numInts2Bit(lc));
            c.paired = paired[i] != 0;
            cands.push_back(std::move(c));
        }
        if (cands.empty()) return out;

        // mismatch ratio filter (include/corrector.hpp:1053-1211)
        {
            std::vector<Cand> kept;
            if (rs.pairedEnd) {
                for (auto& c : cands) {
                    const float ratio = float(c.al.nOps) / float(c.al.overlap);
                    if (c.paired || ratio < opt.pairedFilterThreshold) kept.push_back(std::move(c));
                }
            } else {
                std::vector<float> ratios;
                for (auto& c : cands) ratios.push_back(float(c.al.nOps) / float(c.al.overlap));
                const float thr = seFilterThreshold(ratios.data(), int(ratios.size()), opt.estimatedErrorrate,
                                                    opt.estimatedCoverage, opt.m_coverage);
                for (auto& c : cands) {
                    const float ratio = float(c.al.nOps) / float(c.al.overlap);
                    if (ratio < thr) kept.push_back(std::move(c));
                }
            }
            cands.swap(kept);
        }
        if (cands.empty()) return out;

        // qualities (include/corrector.hpp:1214-1240): the whole pitch-wide
        // row is reversed for reverse-complement candidates
        std::string anchorQ;
        if (opt.useQualityScores) {
            std::vector<char> row(pitchQ, 0);
            rs.gatherQuality(row.data(), id, pitchQ);
            anchorQ.assign(row.begin(), row.end());
            for (auto& c : cands) {
                std::vector<char> q(pitchQ, 0);
                rs.gatherQuality(q.data(), c.id, pitchQ);
                if (c.flag == 2) std::reverse(q.begin(), q.end());
                c.qual.assign(q.begin(), q.end());
            }
        }
        for (auto& c : cands) {
            c.chars.assign(c.len, 'A');
            decode2Bit(&c.chars[0], c.enc.data(), c.len);
            c.weight = overlapWeight(c.al.nOps, c.al.overlap, opt.maxErrorRate);
        }

        MSA msa;
        const int cpitch = rs.maxLen;
        std::vector<char> candChars, candQ;
        std::vector<int> lens, shifts;
        std::vector<float> weights;
        auto buildMsa = [&]() {
            const int n = int(cands.size());
            candChars.assign(std::size_t(n) * cpitch, 0);
            candQ.assign(opt.useQualityScores ? std::size_t(n) * pitchQ : 0, 0);
            lens.resize(n);
            shifts.resize(n);
            weights.resize(n);
            for (int i = 0; i < n; i++) {
                std::memcpy(&candChars[std::size_t(i) * cpitch], cands[i].chars.data(), cands[i].len);
                if (opt.useQualityScores) std::memcpy(&candQ[std::size_t(i) * pitchQ], cands[i].qual.data(), pitchQ);
                lens[i] = cands[i].len;
                shifts[i] = cands[i].al.shift;
                weights[i] = cands[i].weight;
            }
            MSA::Input in{opt.useQualityScores, la, n, cpitch, pitchQ, anchor.data(), candChars.data(),
                          opt.useQualityScores ? anchorQ.data() : nullptr, opt.useQualityScores ? candQ.data() : nullptr,
                          lens.data(), shifts.data(), weights.data()};
            msa.build(in);
        };
        buildMsa();

        // refinement (include/corrector.hpp:1307-1421); none in the gated mode
        for (int it = 0; it < (opt.correctionType == 3 && !opt.gateRefine ? 0 : opt.maxRefinements); it++) {
            std::vector<char> remove;
            if (!msa.differentRegion(int(opt.estimatedCoverage), remove)) break;
            std::vector<Cand> kept;
            for (std::size_t i = 0; i < cands.size(); i++)
                if (!remove[i]) kept.push_back(std::move(cands[i]));
            cands.swap(kept);
            buildMsa();
        }

        // anchor correction (include/corrector.hpp:1423-1472)
        const MSA::Props props = msa.properties(opt.estimatedErrorrate, opt.estimatedCoverage, opt.m_coverage);
        std::string corrected;
        bool isCorrected = false;
        if (opt.correctionType == 1) {
            // correctAnchorClf: consensus, non-HQ columns that differ from
            // the anchor revert unless the classifier accepts
            corrected.assign(msa.consensus.begin() + msa.anchorBegin, msa.consensus.begin() + msa.anchorBegin + la);
            if (!props.isHQ) {
                for (int i = 0; i < la; i++) {
                    if (anchor[i] != msa.consensus[msa.anchorBegin + i]) {
                        float f[13];
                        anchorFeatures(f, msa, props, i);
                        if (!forestA->decide(f)) corrected[i] = anchor[i];
                    }
                }
            }
            isCorrected = true;
        } else if (opt.correctionType == 3) {
            // the engine's gated mode: consensus, and in a non-HQ MSA every
            // column that differs from the anchor is left to the gate
            corrected.assign(msa.consensus.begin() + msa.anchorBegin, msa.consensus.begin() + msa.anchorBegin + la);
            if (!props.isHQ) {
                std::vector<int> dec;
                for (int i = 0; i < la; i++)
                    if (anchor[i] != msa.consensus[msa.anchorBegin + i]) dec.push_back(i);
                for (int i : dec) {
                    GateColumn g;
                    g.id = id;
                    g.pos = std::uint16_t(i);
                    g.orig = std::uint8_t(baseCode(anchor[i]));
                    g.cons = std::uint8_t(baseCode(msa.consensus[msa.anchorBegin + i]));
                    gateFeatures(g.f, msa, props, anchor, anchorQ, msa.anchorBegin, i, int(dec.size()), int(cands.size()));
                    out.gate.push_back(g);
                    if (opt.gateConstant < 0) corrected[i] = anchor[i];
                }
            }
            isCorrected = true;
        } else if (msa.nCandidates > 0) {
            corrected = msa.correctedAnchor(props, opt.estimatedErrorrate, opt.estimatedCoverage, opt.m_coverage);
            isCorrected = true;
        }
        out.hq = props.isHQ && isCorrected;
        if (dump) {
            for (auto& c : cands) dump->msaIds.push_back(c.id);
            dump->consensus.assign(msa.consensus.begin(), msa.consensus.end());
            dump->anchorBegin = msa.anchorBegin;
            dump->corrected = isCorrected;
            dump->hq = props.isHQ;
            dump->correctedSeq = corrected;
        }
        if (isCorrected) {
            out.hasAnchor = true;
            out.anchor.readId = id;
            out.anchor.isAnchor = true;
            out.anchor.hq = props.isHQ;
            makeEdits(out.anchor, corrected, anchor.data(), rs.ambig[id] != 0);
        }

        // candidate correction: classic (src/msa.cpp:457-522) takes candidates
        // starting within +-ncc columns of the anchor and ending at most ncc
        // columns after it; forest (include/corrector.hpp:1539-1623) drops
        // the bound on the start after the anchor start
        if (props.isHQ && opt.correctCandidates) {
            for (int c = 0; c < msa.nCandidates; c++) {
                const int qb = msa.anchorBegin + cands[c].al.shift;
                const int len = cands[c].len;
                const int qe = qb + len;
                const int ncc = opt.new_columns_to_correct;
                const bool forest = opt.correctionTypeCands == 1;
                // the engine's learned gate on candidate columns (gated mode):
                // the forest's range and rule, the gate deciding
                const bool gateC = opt.correctionType == 3 && opt.correctionTypeCands == 3;
                const bool inRange = candidateInRange(qb, qe, msa.anchorBegin, msa.anchorEnd, ncc, forest || gateC);
                if (!inRange) continue;
                std::string seq(msa.consensus.begin() + qb, msa.consensus.begin() + qe);
                if (forest) {
                    const MSA::Props cprops = msa.properties(qb, qe, opt.estimatedErrorrate, opt.estimatedCoverage, opt.m_coverage);
                    for (int i = 0; i < len; i++) {
                        if (cands[c].chars[i] != msa.consensus[qb + i]) {
                            float f[14];
                            candFeatures(f, msa, cprops, cands[c].chars[i], qb, qe, i);
                            if (!forestC->decide(f)) seq[i] = cands[c].chars[i];
                        }
                    }
                } else if (gateC) {
                    const MSA::Props cprops = msa.properties(qb, qe, opt.estimatedErrorrate, opt.estimatedCoverage, opt.m_coverage);
                    std::vector<int> dec;
                    for (int i = 0; i < len; i++)
                        if (cands[c].chars[i] != msa.consensus[qb + i]) dec.push_back(i);
                    const std::string noQ(len, '\0');
                    const bool rcC = cands[c].flag == 2;
                    for (int i : dec) {
                        GateColumn g;
                        g.id = cands[c].id;
                        // forward-strand position and bases of the candidate
                        g.pos = std::uint16_t(rcC ? len - 1 - i : i);
                        const int o = baseCode(cands[c].chars[i]), cn = baseCode(msa.consensus[qb + i]);
                        g.orig = std::uint8_t(rcC ? 3 - o : o);
                        g.cons = std::uint8_t(rcC ? 3 - cn : cn);
                        gateFeatures(g.f, msa, cprops, cands[c].chars, opt.useQualityScores ? cands[c].qual : noQ, qb, i,
                                     int(dec.size()), int(cands.size()));
                        out.gateCands.push_back(g);
                        if (opt.gateConstantCands < 0) seq[i] = cands[c].chars[i];
                    }
                }
                Result r;
                r.readId = cands[c].id;
                r.isAnchor = false;
                std::string fwdCorr(len, 'A'), fwdOrig(len, 'A');
                if (cands[c].flag == 1) {
                    fwdCorr = seq;
                    fwdOrig = cands[c].chars;
                } else {
                    revcompChars(&fwdCorr[0], seq.data(), len);
                    revcompChars(&fwdOrig[0], cands[c].chars.data(), len);
                }
                makeEdits(r, fwdCorr, fwdOrig.data(), rs.ambig[cands[c].id] != 0);
                out.cands.push_back(std::move(r));
            }
        }
        return out;
    }

    float weightOf(const MSA& msa, char b, int c) const {
        switch (b) {
            case 'A': return msa.wA[c];
            case 'C': return msa.wC[c];
            case 'G': return msa.wG[c];
            case 'T': return msa.wT[c];
            default: return 0.f;
        }
    }
    float countOf(const MSA& msa, char b, int c) const {
        switch (b) {
            case 'A': return float(msa.cA[c]);
            case 'C': return float(msa.cC[c]);
            case 'G': return float(msa.cG[c]);
            case 'T': return float(msa.cT[c]);
            default: return 0.f;
        }
    }

    // extract_anchor_v3 (include/classification.hpp:748-800)
    void anchorFeatures(float* f, const MSA& msa, const MSA::Props& p, int i) const {
        const int pos = msa.anchorBegin + i;
        const float cov = opt.estimatedCoverage;
        const float wsum = msa.wA[pos] + msa.wC[pos] + msa.wG[pos] + msa.wT[pos];
        const float wCons = weightOf(msa, msa.consensus[pos], pos);
        const float cCons = countOf(msa, msa.consensus[pos], pos);
        const int coverage = msa.coverage[pos];
        f[0] = float(msa.origCoverages[pos]) / float(coverage);
        f[1] = msa.origWeights[pos] / wsum;
        f[2] = msa.origWeights[pos] / float(msa.origCoverages[pos]);
        f[3] = cCons / float(coverage);
        f[4] = msa.support[pos];
        f[5] = wCons / cCons;
        f[6] = wsum / float(coverage);
        f[7] = float(coverage) / cov;
        f[8] = wsum / cov;
        f[9] = p.avg_support;
        f[10] = p.min_support;
        f[11] = float(p.max_coverage) / cov;
        f[12] = float(p.min_coverage) / cov;
    }

    // extract_cands_v3 (include/classification.hpp:803-877); p are the MSA
    // properties over the candidate's columns [qb, qe)
    void candFeatures(float* f, const MSA& msa, const MSA::Props& p, char origBase, int qb, int qe, int i) const {
        const int pos = qb + i;
        const float cov = opt.estimatedCoverage;
        const float wsum = msa.wA[pos] + msa.wC[pos] + msa.wG[pos] + msa.wT[pos];
        const float wCons = weightOf(msa, msa.consensus[pos], pos);
        const float cCons = countOf(msa, msa.consensus[pos], pos);
        const float wOrig = weightOf(msa, origBase, pos);
        const float cOrig = countOf(msa, origBase, pos);
        const int coverage = msa.coverage[pos];
        f[0] = cOrig / float(coverage);
        f[1] = wOrig / wsum;
        f[2] = wOrig / cOrig;
        f[3] = cCons / float(coverage);
        f[4] = msa.support[pos];
        f[5] = wCons / cCons;
        f[6] = wsum / float(coverage);
        f[7] = float(coverage) / cov;
        f[8] = wsum / cov;
        f[9] = p.avg_support;
        f[10] = p.min_support;
        f[11] = float(p.max_coverage) / cov;
        f[12] = float(p.min_coverage) / cov;
        const int ab = msa.anchorBegin, ae = msa.anchorEnd;
        f[13] = float(std::min(ae, qe) - std::max(ab, qb)) / float(std::max(ae, qe) - std::min(ab, qb));
    }

    // Corrects anchors [0, numAnchors); SE one at a time, PE in mate pairs.
    void run(std::size_t numAnchors, std::vector<Output>& outs, std::vector<AnchorDump>* dumps) const {
        outs.assign(numAnchors, Output{});
        if (dumps) dumps->assign(numAnchors, AnchorDump{});
        const std::size_t step = rs.pairedEnd ? 2 : 1;
        const std::size_t units = (numAnchors + step - 1) / step;
#pragma omp parallel for schedule(dynamic, 64)
        for (std::size_t u = 0; u < units; u++) {
            const std::size_t a0 = u * step;
            const std::size_t a1 = std::min(numAnchors, a0 + step);
            std::vector<std::vector<u32>> lists;
            for (std::size_t a = a0; a < a1; a++) lists.push_back(candidatesOf(u32(a)));
            std::vector<std::vector<char>> paired(lists.size());
            for (std::size_t j = 0; j < lists.size(); j++) paired[j].assign(lists[j].size(), 0);
            // include/corrector.hpp:763-807 with include/util.hpp:1021-1199:
            // candidate x of one mate is paired iff its mate id x^1 is a
            // candidate of the other mate
            if (rs.pairedEnd && lists.size() == 2) {
                for (int j = 0; j < 2; j++) {
                    const auto& other = lists[1 - j];
                    for (std::size_t i = 0; i < lists[j].size(); i++)
                        paired[j][i] = std::binary_search(other.begin(), other.end(), lists[j][i] ^ 1u);
                }
            }
            for (std::size_t j = 0; j < lists.size(); j++) {
                const std::size_t a = a0 + j;
                outs[a] = correctOne(u32(a), lists[j], paired[j], dumps ? &(*dumps)[a] : nullptr);
            }
        }
    }
};

// Groups results per read and applies the combine rule; returns the
// corrected sequence of every read (or the original) and a label code.
struct Combined {
    bool corrected = false;
    int label = 0;  // 1 lq anchor only, 2 lq with candidates, 3 hq
};

Combined combine(std::vector<const Result*>& res, std::string& seq) {
    Combined cr;
    if (res.empty()) return cr;
    const Result* anchor = nullptr;
    for (auto* r : res)
        if (r->isAnchor) { anchor = r; break; }
    if (!anchor) return cr;
    auto apply = [&](const Result* r) {
        if (r->useEdits) {
            for (auto& e : r->edits) seq[e.first] = e.second;
        } else {
            seq = r->sequence;
        }
    };
    if (anchor->hq) {
        apply(anchor);
        cr.corrected = true;
        cr.label = 3;
        return cr;
    }
    if (res.size() >= 3) {
        // all results must yield the same sequence as the first one
        auto materialize = [&](const Result* r) {
            std::string s = seq;
            if (r->useEdits) {
                for (auto& e : r->edits) s[e.first] = e.second;
            } else {
                s = r->sequence;
            }
            return s;
        };
        const std::string first = materialize(res[0]);
        for (std::size_t i = 1; i < res.size(); i++) {
            const Result* r = res[i];
            bool eq;
            if (res[0]->useEdits && r->useEdits) eq = res[0]->edits == r->edits;
            else if (!res[0]->useEdits && !r->useEdits) eq = res[0]->sequence == r->sequence;
            else eq = materialize(r) == first && (res[0]->useEdits ? materialize(res[0]) : res[0]->sequence).size() == seq.size();
            if (!eq) return cr;
        }
        apply(res[0]);
        cr.corrected = true;
        cr.label = 2;
        return cr;
    }
    apply(anchor);
    cr.corrected = true;
    cr.label = 1;
    return cr;
}

// ---------------------------------------------------------------------------
// CLI (subset of src/options.cpp:705-895)
// ---------------------------------------------------------------------------
std::size_t parseMem(const std::string& s) {
    double v = std::stod(s);
    const char u = s.empty() ? 0 : s.back();
    if (u == 'K' || u == 'k') v *= 1024.0;
    if (u == 'M' || u == 'm') v *= 1024.0 * 1024.0;
    if (u == 'G' || u == 'g') v *= 1024.0 * 1024.0 * 1024.0;
    return std::size_t(v);
}

Options parseArgs(int argc, char** argv) {
    Options o;
    auto isFlagTrue = [](const std::string& v) { return v != "false" && v != "0"; };
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string val;
        const auto eq = a.find('=');
        bool hasEq = false;
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
            hasEq = true;
        }
        auto next = [&]() -> std::string {
            if (hasEq) return val;
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
            return argv[++i];
        };
        auto boolFlag = [&]() -> bool {
            if (hasEq) return isFlagTrue(val);
            return true;
        };
        if (a == "-i" || a == "--inputfiles") o.inputfiles.push_back(next());
        else if (a == "-o" || a == "--outputfilenames") o.outputfilenames.push_back(next());
        else if (a == "-d" || a == "--outdir") o.outputdirectory = next();
        else if (a == "-c" || a == "--coverage") o.estimatedCoverage = std::stof(next());
        else if (a == "--pairmode") {
            std::string m = next();
            o.pairedEnd = (m == "PE" || m == "pe");
        } else if (a == "-k" || a == "--kmerlength") o.kmerlength = std::stoi(next());
        else if (a == "-h" || a == "--hashmaps") o.numHashFunctions = std::stoi(next());
        else if (a == "-t" || a == "--threads") o.threads = std::stoi(next());
        else if (a == "-q" || a == "--useQualityScores") o.useQualityScores = boolFlag();
        else if (a == "--qualityScoreBits") o.qualityScoreBits = std::stoi(next());
        else if (a == "--excludeAmbiguous") o.excludeAmbiguousReads = boolFlag();
        else if (a == "--maxmismatchratio") o.maxErrorRate = std::stof(next());
        else if (a == "--minalignmentoverlap") o.min_overlap = std::stoi(next());
        else if (a == "--minalignmentoverlapratio") o.min_overlap_ratio = std::stof(next());
        else if (a == "--errorfactortuning") o.estimatedErrorrate = std::stof(next());
        else if (a == "--coveragefactortuning") o.m_coverage = std::stof(next());
        else if (a == "--candidateCorrection") o.correctCandidates = boolFlag();
        else if (a == "--candidateCorrectionNewColumns") o.new_columns_to_correct = std::stoi(next());
        else if (a == "--correctionQualityLabels") o.outputCorrectionQualityLabels = boolFlag();
        else if (a == "--correctionType") o.correctionType = std::stoi(next());
        else if (a == "--correctionTypeCands") o.correctionTypeCands = std::stoi(next());
        else if (a == "--candidateLimit") o.candidateLimit = std::stoi(next());
        else if (a == "--refinements") o.maxRefinements = std::stoi(next());
        else if (a == "--gateRefine") o.gateRefine = true;
        else if (a == "--gateConstant") o.gateConstant = std::stoi(next());
        else if (a == "--gateDump") o.gateDump = next();
        else if (a == "--gateConstantCands") o.gateConstantCands = std::stoi(next());
        else if (a == "--gateDumpCands") o.gateDumpCands = next();
        else if (a == "--ml-forestfile") o.forestAnchor = next();
        else if (a == "--ml-cands-forestfile") o.forestCands = next();
        else if (a == "--thresholdAnchor") o.thresholdAnchor = std::stof(next());
        else if (a == "--thresholdCands") o.thresholdCands = std::stof(next());
        else if (a == "--pairedFilterThreshold") o.pairedFilterThreshold = std::stof(next());
        else if (a == "--fixedNumberOfReads") o.fixedNumberOfReads = std::stoull(next());
        else if (a == "--dump") o.dumpFile = next();
        else if (a == "--maxForestTreesAnchor") {
            const int v = std::stoi(next());
            o.maxTreesAnchor = v > 0 ? std::uint32_t(v) : 0xFFFFFFFFu;
        } else if (a == "--maxForestTreesCands") {
            const int v = std::stoi(next());
            o.maxTreesCands = v > 0 ? std::uint32_t(v) : 0xFFFFFFFFu;
        }
        else if (a == "-m" || a == "--memTotal" || a == "--memHashtables" || a == "--tempdir" || a == "--hashloadfactor" ||
                 a == "-g" || a == "--gpu" || a == "--batchsize" || a == "--warpcore")
            (void)next();
        else if (a == "-p" || a == "--showProgress" || a == "--gzoutput" || a == "--enforceHashmapCount") (void)0;
        else throw std::runtime_error("unknown option " + a);
    }
    if (o.inputfiles.empty() || o.outputfilenames.empty()) throw std::runtime_error("need -i and -o");
    return o;
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int runMain(int argc, char** argv) {
    Options opt = parseArgs(argc, argv);
    omp_set_num_threads(std::max(1, opt.threads));
    const double t0 = now();
    std::vector<std::size_t> fileStarts;
    std::vector<Read> reads;
    if (opt.pairedEnd && opt.inputfiles.size() == 2) {
        reads = readAll(opt);
        fileStarts.push_back(0);
    } else {
        for (const auto& f : opt.inputfiles) {
            fileStarts.push_back(reads.size());
            Options one = opt;
            one.inputfiles = {f};
            one.pairedEnd = false;
            auto part = readAll(one);
            for (auto& r : part) reads.push_back(std::move(r));
        }
    }
    std::vector<std::string> original(reads.size());
    for (std::size_t i = 0; i < reads.size(); i++) original[i] = reads[i].seq;
    ReadStore rs = buildReadStore(reads, opt, fileStarts);
    Tables tables;
    tables.build(rs, opt.kmerlength, opt.numHashFunctions, resultsPerMapThreshold(opt.estimatedCoverage));
    const double t1 = now();

    Corrector corr(opt, rs, tables);
    // include/classification.hpp:72-74, descriptors of the v3 extractors
    Forest fa, fc;
    if (opt.correctionType == 1) {
        fa.load(opt.forestAnchor, "13 extract_anchor_v3", opt.maxTreesAnchor, opt.thresholdAnchor);
        corr.forestA = &fa;
    }
    if (opt.correctionTypeCands == 1) {
        fc.load(opt.forestCands, "14 extract_cands_v3", opt.maxTreesCands, opt.thresholdCands);
        corr.forestC = &fc;
    }
    std::size_t numAnchors = opt.fixedNumberOfReads ? std::min(opt.fixedNumberOfReads, rs.n) : rs.n;
    std::vector<Corrector::Output> outs;
    std::vector<AnchorDump> dumps;
    corr.run(numAnchors, outs, opt.dumpFile.empty() ? nullptr : &dumps);
    const double t2 = now();
    // per decision column: u32 read id, u16 position, u8 original and
    // consensus base, 32 float features; anchors in order
    auto dumpGate = [&](const std::string& path, bool cands) {
        std::ofstream g(path, std::ios::binary);
        for (std::size_t a = 0; a < numAnchors; a++)
            for (const auto& c : cands ? outs[a].gateCands : outs[a].gate) {
                g.write(reinterpret_cast<const char*>(&c.id), 4);
                g.write(reinterpret_cast<const char*>(&c.pos), 2);
                g.write(reinterpret_cast<const char*>(&c.orig), 1);
                g.write(reinterpret_cast<const char*>(&c.cons), 1);
                g.write(reinterpret_cast<const char*>(c.f), sizeof(c.f));
            }
    };
    if (!opt.gateDump.empty()) dumpGate(opt.gateDump, false);
    if (!opt.gateDumpCands.empty()) dumpGate(opt.gateDumpCands, true);

    // results per read; HQ anchors with zero edits are not stored and
    // candidate results of reads corrected as HQ anchors are dropped
    // (src/correct_cpu.cpp:262-266, include/corrector.hpp:1693-1697)
    std::vector<char> hq(rs.n, 0);
    for (std::size_t a = 0; a < numAnchors; a++)
        if (outs[a].hasAnchor && outs[a].anchor.hq) hq[a] = 1;
    std::vector<std::vector<const Result*>> per(rs.n);
    for (std::size_t a = 0; a < numAnchors; a++) {
        const auto& o = outs[a];
        if (o.hasAnchor) {
            const Result& r = o.anchor;
            if (!(r.hq && r.useEdits && r.edits.empty())) per[a].push_back(&r);
        }
        for (const auto& c : o.cands)
            if (!hq[c.readId]) per[c.readId].push_back(&c);
    }
    // candidate results first appear in ascending anchor order; the anchor
    // entry is located by search, so order within a read does not matter
    long long lastWithResult = -1;
    for (std::size_t i = 0; i < rs.n; i++)
        if (!per[i].empty()) lastWithResult = (long long)i;

    const std::string outPath = opt.outputdirectory.empty() ? opt.outputfilenames[0]
                                                            : opt.outputdirectory + "/" + opt.outputfilenames[0];
    const bool fastq = isFastq(opt.inputfiles[0]);
    if (!opt.outputdirectory.empty()) mkdir(opt.outputdirectory.c_str(), 0755);
    std::ofstream os(outPath);
    if (!os) throw std::runtime_error("cannot open output " + outPath);
    std::string buf;
    for (std::size_t i = 0; i < rs.n; i++) {
        std::string seq = original[i];
        Combined cr = combine(per[i], seq);
        buf += fastq ? '@' : '>';
        buf += reads[i].header;
        if (opt.outputCorrectionQualityLabels && (long long)i <= lastWithResult) {
            buf += cr.corrected ? (cr.label == 1 ? " care:q=1" : cr.label == 2 ? " care:q=2" : " care:q=3") : " care:q=0";
        }
        buf += '\n';
        buf += seq;
        buf += '\n';
        if (fastq) {
            buf += "+\n";
            buf += reads[i].qual;
            buf += '\n';
        }
        if (buf.size() > (1u << 22)) {
            os.write(buf.data(), buf.size());
            buf.clear();
        }
    }
    os.write(buf.data(), buf.size());
    os.close();

    if (!opt.dumpFile.empty()) {
        std::ofstream d(opt.dumpFile);
        for (const auto& x : dumps) {
            d << "anchor " << x.id << " ncand " << x.candidates.size();
            d << " ids";
            for (auto v : x.candidates) d << ' ' << v;
            d << " flags";
            for (auto v : x.flags) d << ' ' << v;
            d << " aln";
            for (auto& v : x.chosen) d << ' ' << v.shift << ':' << v.overlap << ':' << v.nOps;
            d << " msa";
            for (auto v : x.msaIds) d << ' ' << v;
            d << " begin " << x.anchorBegin << " cons " << (x.consensus.empty() ? "-" : x.consensus);
            d << " corr " << int(x.corrected) << " hq " << int(x.hq) << " seq " << (x.correctedSeq.empty() ? "-" : x.correctedSeq) << '\n';
        }
    }
    const double t3 = now();
    std::fprintf(stderr, "care-oracle: reads %zu anchors %zu step1 %.3f s step2 %.3f s step3 %.3f s -> %.1f reads/s\n", rs.n,
                 numAnchors, t1 - t0, t2 - t1, t3 - t2, double(numAnchors) / (t2 - t1));
    return 0;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface for the tests
// ---------------------------------------------------------------------------
extern "C" {

int orc_main(int argc, char** argv) {
    try {
        return orc::runMain(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "care-oracle: %s\n", e.what());
        return 1;
    }
}

void orc_encode_2bit(const char* seq, int len, std::uint32_t* out) { orc::encode2Bit(out, seq, len); }
void orc_decode_2bit(const std::uint32_t* enc, int len, char* out) { orc::decode2Bit(out, enc, len); }
void orc_revcomp_2bit(const std::uint32_t* enc, int len, std::uint32_t* out) { orc::revcomp2Bit(out, enc, len); }
void orc_to_hilo(const std::uint32_t* enc, int len, std::uint32_t* out) { orc::toHiLo(out, enc, len); }
void orc_encode_quality(const char* q, int len, int bits, std::uint32_t* out) { orc::encodeQuality(out, q, len, bits); }
void orc_decode_quality(const std::uint32_t* enc, int len, int bits, char* out) { orc::decodeQuality(out, enc, len, bits); }
std::uint64_t orc_murmur64(std::uint64_t x) { return orc::murmur64(x); }
void orc_signature(const std::uint32_t* enc, int len, int k, int h, std::uint64_t* out) { orc::signature(out, enc, len, k, h); }
float orc_quality_weight(int c) { return orc::qweights().w[c & 255]; }

// returns valid; writes score, overlap, shift, nOps
int orc_align(const std::uint32_t* a, int la, const std::uint32_t* c, int lc, int minOverlap, float maxErrorRate,
              float minOverlapRatio, int* out4) {
    const auto r = orc::align(a, la, c, lc, minOverlap, maxErrorRate, minOverlapRatio);
    out4[0] = r.score;
    out4[1] = r.overlap;
    out4[2] = r.shift;
    out4[3] = r.nOps;
    return r.valid ? 1 : 0;
}

int orc_choose_orientation(const int* f4, int fvalid, const int* r4, int rvalid, int la, float minOverlapRatio, int minOverlap,
                           float estimatedErrorrate) {
    orc::Alignment f, r;
    f.score = f4[0]; f.overlap = f4[1]; f.shift = f4[2]; f.nOps = f4[3]; f.valid = fvalid != 0;
    r.score = r4[0]; r.overlap = r4[1]; r.shift = r4[2]; r.nOps = r4[3]; r.valid = rvalid != 0;
    return orc::chooseOrientation(f, r, la, minOverlapRatio, minOverlap, estimatedErrorrate * 4.0f);
}

// Minhash tables over a packed read set; query returns sorted, unique
// candidate ids without the anchor itself.
void* orc_tables_build(const std::uint32_t* seqs, int pitch, const int* lens, std::size_t n, int k, int h, float coverage) {
    auto* rs = new orc::ReadStore();
    rs->n = n;
    rs->pitch = pitch;
    rs->seqs.assign(seqs, seqs + n * std::size_t(pitch));
    rs->lens.assign(lens, lens + n);
    rs->ambig.assign(n, 0);
    auto* t = new orc::Tables();
    t->build(*rs, k, h, orc::resultsPerMapThreshold(coverage));
    delete rs;
    return t;
}
void orc_tables_free(void* t) { delete static_cast<orc::Tables*>(t); }
int orc_tables_query(void* t, const std::uint32_t* enc, int len, std::uint32_t self, std::uint32_t* out, int cap) {
    std::vector<std::uint32_t> ids;
    static_cast<orc::Tables*>(t)->query(ids, enc, len);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    ids.erase(std::remove(ids.begin(), ids.end(), self), ids.end());
    const int n = int(ids.size());
    for (int i = 0; i < n && i < cap; i++) out[i] = ids[i];
    return n;
}
// Candidate limit of the engine's gated fast path (no counterpart in the
// reference): of the de-duplicated candidates keep the `limit` ones hit by
// the most maps, ties going to the smaller read id, returned in ascending id
// order.
int orc_tables_query_limited(void* t, const std::uint32_t* enc, int len, std::uint32_t self, int limit,
                             std::uint32_t* out, int cap) {
    std::vector<std::uint32_t> raw;
    static_cast<orc::Tables*>(t)->query(raw, enc, len);
    std::sort(raw.begin(), raw.end());
    std::vector<std::pair<int, std::uint32_t>> hc;  // (hits, id)
    for (std::size_t i = 0; i < raw.size();) {
        std::size_t j = i;
        while (j < raw.size() && raw[j] == raw[i]) j++;
        if (raw[i] != self) hc.push_back({int(j - i), raw[i]});
        i = j;
    }
    if (limit > 0 && int(hc.size()) > limit) {
        std::stable_sort(hc.begin(), hc.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        hc.resize(std::size_t(limit));
        std::sort(hc.begin(), hc.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
    }
    const int n = int(hc.size());
    for (int i = 0; i < n && i < cap; i++) out[i] = hc[std::size_t(i)].second;
    return n;
}
// ---- single rules, for the known-answer tests ----

namespace {
orc::MSA msaFrom(int useQ, int la, const char* anchor, const char* anchorQ, int n, const char* cands, int candPitch,
                 const char* candQ, int qualPitch, const int* lens, const int* shifts, const float* weights) {
    orc::MSA msa;
    orc::MSA::Input in{useQ != 0, la, n, candPitch, qualPitch, anchor, cands, anchorQ, candQ, lens, shifts, weights};
    msa.build(in);
    return msa;
}
}  // namespace

// MSA of an anchor and candidates given as characters (src/msa.cpp:9-173):
// dims = {nColumns, anchorBegin}; column arrays hold cap columns (counts and
// weights [4][cap] in A, C, G, T order).  Returns -1 if cap is too small.
int orc_msa_build(int useQ, int la, const char* anchor, const char* anchorQ, int n, const char* cands, int candPitch,
                  const char* candQ, int qualPitch, const int* lens, const int* shifts, const float* weights, int cap,
                  int* dims, char* consensus, float* support, int* coverage, int* counts, float* wts, int* origCov) {
    const orc::MSA m = msaFrom(useQ, la, anchor, anchorQ, n, cands, candPitch, candQ, qualPitch, lens, shifts, weights);
    dims[0] = m.nColumns;
    dims[1] = m.anchorBegin;
    if (m.nColumns > cap) return -1;
    for (int c = 0; c < m.nColumns; c++) {
        consensus[c] = m.consensus[c];
        support[c] = m.support[c];
        coverage[c] = m.coverage[c];
        counts[0 * cap + c] = m.cA[c];
        counts[1 * cap + c] = m.cC[c];
        counts[2 * cap + c] = m.cG[c];
        counts[3 * cap + c] = m.cT[c];
        wts[0 * cap + c] = m.wA[c];
        wts[1 * cap + c] = m.wC[c];
        wts[2 * cap + c] = m.wG[c];
        wts[3 * cap + c] = m.wT[c];
        origCov[c] = m.origCoverages[c];
    }
    return 0;
}

// MSA properties over columns [first, last) of the given support and
// coverage (src/msa.cpp:304-376): out4 = {avg support, min support, min
// coverage, max coverage}; returns the high-quality flag.
int orc_msa_properties(const float* support, const int* coverage, int ncols, int first, int last, float err, float cov,
                       float mcov, float* out4) {
    orc::MSA m;
    m.nColumns = ncols;
    m.support.assign(support, support + ncols);
    m.coverage.assign(coverage, coverage + ncols);
    const auto p = m.properties(first, last, err, cov, mcov);
    out4[0] = p.avg_support;
    out4[1] = p.min_support;
    out4[2] = float(p.min_coverage);
    out4[3] = float(p.max_coverage);
    return p.isHQ ? 1 : 0;
}

// Classic anchor correction of the MSA built from the arguments
// (src/msa.cpp:379-453); returns the high-quality flag.
int orc_msa_corrected_anchor(int useQ, int la, const char* anchor, const char* anchorQ, int n, const char* cands,
                             int candPitch, const char* candQ, int qualPitch, const int* lens, const int* shifts,
                             const float* weights, float err, float cov, float mcov, char* out) {
    const orc::MSA m = msaFrom(useQ, la, anchor, anchorQ, n, cands, candPitch, candQ, qualPitch, lens, shifts, weights);
    const auto p = m.properties(err, cov, mcov);
    const std::string c = m.correctedAnchor(p, err, cov, mcov);
    std::memcpy(out, c.data(), std::size_t(la));
    return p.isHQ ? 1 : 0;
}

// Different-region search of the MSA built from the arguments
// (src/msa.cpp:532-734): returns 1 if a column was found, with the
// candidates to remove flagged in remove[n].
int orc_msa_different_region(int la, const char* anchor, int n, const char* cands, int candPitch, const int* lens,
                             const int* shifts, const float* weights, int datasetCoverage, char* remove) {
    const orc::MSA m = msaFrom(0, la, anchor, nullptr, n, cands, candPitch, nullptr, 0, lens, shifts, weights);
    std::vector<char> rem;
    const bool found = m.differentRegion(datasetCoverage, rem);
    for (int i = 0; i < n; i++) remove[i] = found && !rem.empty() ? rem[std::size_t(i)] : 0;
    return found ? 1 : 0;
}

float orc_se_filter_threshold(const float* ratios, int n, float err, float cov, float mcov) {
    return orc::seFilterThreshold(ratios, n, err, cov, mcov);
}

int orc_candidate_in_range(int qb, int qe, int ab, int ae, int ncc, int forestOrGate) {
    return orc::candidateInRange(qb, qe, ab, ae, ncc, forestOrGate != 0) ? 1 : 0;
}

float orc_overlap_weight(int nOps, int overlap, float maxMismatchRatio) {
    return orc::overlapWeight(nOps, overlap, maxMismatchRatio);
}

// Edit list of a corrected read (include/corrector.hpp:1639-1761): the
// number of edits (positions and bases written) if it is stored as edits,
// -1 if it is stored as the whole sequence.
int orc_make_edits(const char* corrected, const char* original, int len, int ambiguous, int* positions, char* bases) {
    orc::Result r;
    orc::makeEdits(r, std::string(corrected, std::size_t(len)), original, ambiguous != 0);
    if (!r.useEdits) return -1;
    for (std::size_t i = 0; i < r.edits.size(); i++) {
        positions[i] = r.edits[i].first;
        bases[i] = r.edits[i].second;
    }
    return int(r.edits.size());
}

// Forest vote (include/forest.hpp:217-226) of the .rf file at path for one
// feature vector; kind 0 anchor (13 features), 1 candidates (14).  Returns
// 1 accept, 0 reject, -1 if the file does not load.
int orc_forest_decide(const char* path, int kind, std::uint32_t maxTrees, float threshold, const float* features) {
    try {
        orc::Forest f;
        f.load(path, kind == 0 ? "13 extract_anchor_v3" : "14 extract_cands_v3", maxTrees, threshold);
        return f.decide(features) ? 1 : 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Combine rule of one read (src/correctionresultoutput.cpp:730-1023): n
// results in stored order, flags bit 0 anchor, bit 1 high quality, each a
// corrected sequence of the read's length (stored as edits or whole as the
// edit rule decides); writes the output sequence, returns the label (0
// uncorrected, 1 low-quality anchor alone, 2 agreeing results, 3 HQ anchor).
int orc_combine(const char* original, int len, int ambiguous, int n, const int* flags, const char* corrected, char* out) {
    std::vector<orc::Result> res(static_cast<std::size_t>(n));
    std::vector<const orc::Result*> ptrs;
    for (int i = 0; i < n; i++) {
        res[i].isAnchor = (flags[i] & 1) != 0;
        res[i].hq = (flags[i] & 2) != 0;
        orc::makeEdits(res[i], std::string(corrected + std::size_t(i) * len, std::size_t(len)), original, ambiguous != 0);
        ptrs.push_back(&res[i]);
    }
    std::string seq(original, std::size_t(len));
    const auto c = orc::combine(ptrs, seq);
    std::memcpy(out, seq.data(), seq.size());
    return c.label;
}

int orc_tables_bucket(void* t, int map, std::uint64_t key, std::uint32_t* out, int cap) {
    auto* tb = static_cast<orc::Tables*>(t);
    auto it = tb->maps[map].index.find(key);
    if (it == tb->maps[map].index.end()) return 0;
    const int n = int(it->second.second);
    for (int i = 0; i < n && i < cap; i++) out[i] = tb->maps[map].values[it->second.first + i];
    return n;
}
}
