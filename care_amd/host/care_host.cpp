// Host side of the MI355X CARE corrector: the `care-gpu` command line.
//
// STEP1  parse inputs, pack reads (2-bit, N rule, binned qualities), upload
//        to HBM and build the minhash tables on the device;
// STEP2  correct all reads in batches on the device, turning every batch's
//        corrected sequences into edit lists on the host;
// STEP3  combine the results per read and write the ordered output.
//
// Mirrors src/gpu/dispatch_care_correct_gpu.cu:522-929 (three timed steps),
// src/options.cpp:705-895 (options), include/chunkedreadstorageconstruction.hpp
// (read packing), src/correctionresultoutput.cpp:31-445,730-1023 (merge and
// combine) and src/readlibraryio.cpp:30-130 (output format).

#include "../../include/care_amd.h"
#include "../../include/care_host.h"
#include "readset.h"

#include <zlib.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <thread>
#include <vector>

// text of the last host error (care_host_last_error), shared with storage_files.cpp
__attribute__((visibility("hidden"))) thread_local std::string g_hostError;

namespace {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Input: FASTA/FASTQ, plain or gzip, parsed like kseq (header = the whole
// line after '@' or '>', multi-line sequences, '+' line skipped).
// ---------------------------------------------------------------------------
class ByteStream {
public:
    explicit ByteStream(const std::string& path) {
        f_ = gzopen(path.c_str(), "rb");
        if (!f_) throw std::runtime_error("cannot open input file " + path);
        gzbuffer(f_, 1 << 20);
        buf_.resize(1 << 22);
    }
    ~ByteStream() {
        if (f_) gzclose(f_);
    }
    int get() {
        if (pos_ >= end_) {
            if (eof_) return -1;
            const int r = gzread(f_, buf_.data(), unsigned(buf_.size()));
            if (r <= 0) {
                eof_ = true;
                return -1;
            }
            pos_ = 0;
            end_ = size_t(r);
        }
        return (unsigned char)buf_[pos_++];
    }
    // appends bytes up to '\n' (dropped); false if nothing was read
    bool line(std::string& s) {
        bool any = false;
        for (;;) {
            if (pos_ >= end_) {
                if (get() < 0) break;
                pos_--;
            }
            const char* b = buf_.data() + pos_;
            const char* e = buf_.data() + end_;
            const char* nl = static_cast<const char*>(std::memchr(b, '\n', size_t(e - b)));
            any = true;
            if (nl) {
                s.append(b, nl);
                pos_ += size_t(nl - b) + 1;
                break;
            }
            s.append(b, e);
            pos_ = end_;
        }
        if (!s.empty() && s.back() == '\r') s.pop_back();
        return any;
    }

private:
    gzFile f_ = nullptr;
    std::vector<char> buf_;
    size_t pos_ = 0, end_ = 0;
    bool eof_ = false;
};

struct RecordReader {
    ByteStream in;
    int pending = 0;
    bool fastq = false;
    explicit RecordReader(const std::string& p) : in(p) {}
    bool next(std::string& header, std::string& seq, std::string& qual) {
        int c = 0;
        if (pending == 0) {
            while ((c = in.get()) >= 0 && c != '>' && c != '@') {}
            if (c < 0) return false;
            pending = c;
        }
        header.clear();
        seq.clear();
        qual.clear();
        if (!in.line(header)) return false;
        while ((c = in.get()) >= 0 && c != '>' && c != '+' && c != '@') {
            if (c == '\n') continue;
            seq.push_back(char(c));
            in.line(seq);
        }
        pending = (c == '>' || c == '@') ? c : 0;
        if (c != '+') return true;
        fastq = true;
        while ((c = in.get()) >= 0 && c != '\n') {}
        while (qual.size() < seq.size())
            if (!in.line(qual)) break;
        pending = 0;
        if (qual.size() != seq.size()) throw std::runtime_error("quality string length differs from sequence length");
        return true;
    }
};

// Growable byte array whose new bytes are not zeroed: arenas filled by
// bulk copies (memcpy, not element-wise) with no initialising pass.
class Arena {
public:
    const char* data() const { return p_.get(); }
    char* data() { return p_.get(); }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    void reserve(size_t c) {
        if (c <= cap_) return;
        std::unique_ptr<char[]> q(new char[c]);
        if (n_) std::memcpy(q.get(), p_.get(), n_);
        p_ = std::move(q);
        cap_ = c;
    }
    void append(const char* b, size_t len) {
        if (n_ + len > cap_) reserve(std::max(n_ + len, cap_ * 2 + 4096));
        if (len) std::memcpy(p_.get() + n_, b, len);
        n_ += len;
    }
    void resizeUninit(size_t m) {
        reserve(m);
        n_ = m;
    }
    bool operator==(const Arena& o) const { return n_ == o.n_ && (n_ == 0 || std::memcmp(data(), o.data(), n_) == 0); }

private:
    std::unique_ptr<char[]> p_;
    size_t n_ = 0, cap_ = 0;
};

// All reads of a library in read-id order, stored in arenas.
struct Library {
    Arena headers, seqs, quals;
    std::vector<u64> hoff, soff;  // n + 1 offsets (quals share soff)
    std::vector<u32> fileOf;      // first read id of every input file, for the N rule batches
    bool fastq = true;
    u64 size() const { return soff.empty() ? 0 : soff.size() - 1; }
    void add(const std::string& h, const std::string& s, const std::string& q, bool keepQual) {
        if (hoff.empty()) {
            hoff.push_back(0);
            soff.push_back(0);
        }
        headers.append(h.data(), h.size());
        hoff.push_back(headers.size());
        seqs.append(s.data(), s.size());
        if (keepQual) quals.append(q.data(), q.size());
        soff.push_back(seqs.size());
    }
    const char* seq(u64 i) const { return seqs.data() + soff[i]; }
    int len(u64 i) const { return int(soff[i + 1] - soff[i]); }
};

// ---------------------------------------------------------------------------
// Parallel parse of four-line FASTQ files (the usual layout).  The file is
// mapped (or, gzip'ed, inflated into memory) and cut into one byte range
// per thread; the threads count their newlines, a prefix sum gives every
// range its first line number, and a record starts at each line whose
// number is a multiple of four.  A range owns the records that start after
// one of its newlines (range 0 also the first one), parses them into its
// own arenas, and the arenas are copied into the library in parallel.  A
// record that is not four lines ('@' header, sequence, '+' line, quality of
// the sequence's length) makes the caller fall back to the sequential
// parser above, which also handles FASTA and multi-line records; both give
// the same library for four-line files.
// ---------------------------------------------------------------------------
struct InputBytes {
    const char* p = nullptr;
    size_t n = 0;
    std::vector<char> inflated;
    void* map = nullptr;
    explicit InputBytes(const std::string& path) {
        const int fd = open(path.c_str(), O_RDONLY);
        if (fd < 0) throw std::runtime_error("cannot open input file " + path);
        struct stat st {};
        fstat(fd, &st);
        unsigned char magic[2] = {0, 0};
        const bool gz = pread(fd, magic, 2, 0) == 2 && magic[0] == 0x1f && magic[1] == 0x8b;
        if (!gz && st.st_size > 0) {
            map = mmap(nullptr, size_t(st.st_size), PROT_READ, MAP_PRIVATE, fd, 0);
            close(fd);
            if (map == MAP_FAILED) throw std::runtime_error("cannot map input file " + path);
            p = static_cast<const char*>(map);
            n = size_t(st.st_size);
            return;
        }
        close(fd);
        gzFile f = gzopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open input file " + path);
        gzbuffer(f, 1 << 20);
        for (;;) {
            const size_t have = inflated.size();
            inflated.resize(have + (size_t(64) << 20));
            const int r = gzread(f, inflated.data() + have, unsigned(size_t(64) << 20));
            if (r < 0) {
                gzclose(f);
                throw std::runtime_error("read error on " + path);
            }
            inflated.resize(have + size_t(r));
            if (r == 0) break;
        }
        gzclose(f);
        p = inflated.data();
        n = inflated.size();
    }
    ~InputBytes() {
        if (map) munmap(map, n);
    }
    InputBytes(const InputBytes&) = delete;
    InputBytes& operator=(const InputBytes&) = delete;
};

struct LibraryPart {
    Arena headers, seqs, quals;
    std::vector<u64> hend, send;  // end offsets of every record
    bool ok = true;
};

// parses the records starting in [b, stop) (a record may run past stop)
void parseRange(const char* b, const char* stop, const char* e, bool keepQual, LibraryPart& part) {
    auto nextLine = [&](const char*& p, const char*& ls, const char*& le) {
        ls = p;
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', size_t(e - p)));
        le = nl ? nl : e;
        p = nl ? nl + 1 : e;
        if (le > ls && le[-1] == '\r') le--;
    };
    const char* p = b;
    // the range's bytes bound every arena
    const size_t bytes = size_t(std::min(stop, e) - b) + 1;
    part.headers.reserve(bytes / 4);
    part.seqs.reserve(bytes / 2);
    if (keepQual) part.quals.reserve(bytes / 2);
    part.hend.reserve(bytes / 64);
    part.send.reserve(bytes / 64);
    while (p < stop && p < e) {
        const char *h0, *h1, *s0, *s1, *x0, *x1, *q0, *q1;
        nextLine(p, h0, h1);
        nextLine(p, s0, s1);
        nextLine(p, x0, x1);
        nextLine(p, q0, q1);
        if (h1 == h0 || *h0 != '@' || x1 == x0 || *x0 != '+' || (q1 - q0) != (s1 - s0)) {
            part.ok = false;
            return;
        }
        part.headers.append(h0 + 1, size_t(h1 - h0 - 1));
        part.hend.push_back(part.headers.size());
        part.seqs.append(s0, size_t(s1 - s0));
        part.send.push_back(part.seqs.size());
        if (keepQual) part.quals.append(q0, size_t(q1 - q0));
    }
}

// false: not a four-line FASTQ file (lib untouched)
bool readFastq4(const std::string& path, bool keepQual, int threads, Library& out) {
    InputBytes in(path);
    if (in.n == 0 || in.p[0] != '@') return false;
    const char* base = in.p;
    const size_t n = in.n;
    const int T = int(std::max<size_t>(1, std::min<size_t>(size_t(std::max(1, threads)), n >> 16)));
    std::vector<size_t> lo(T + 1);
    for (int t = 0; t <= T; t++) lo[t] = n * size_t(t) / size_t(T);
    auto parallel = [&](auto&& fn) {
        std::vector<std::thread> pool;
        for (int t = 1; t < T; t++) pool.emplace_back(fn, t);
        fn(0);
        for (auto& th : pool) th.join();
    };
    // newlines per range, then the first line number of every range
    std::vector<u64> nl(T + 1, 0);
    parallel([&](int t) { nl[t + 1] = u64(std::count(base + lo[t], base + lo[t + 1], '\n')); });
    for (int t = 0; t < T; t++) nl[t + 1] += nl[t];
    // first record start owned by each range
    std::vector<const char*> first(T + 1, base + n);
    parallel([&](int t) {
        if (t == 0) {
            first[0] = base;
            return;
        }
        u64 k = nl[t];  // newlines before this range
        for (const char* q = base + lo[t]; q < base + lo[t + 1]; q++) {
            if (*q != '\n') continue;
            k++;  // the line after this newline has number k
            if (k % 4 == 0) {
                first[t] = q + 1;
                return;
            }
        }
    });
    // a range parses from its first record to the next range's first record
    for (int t = T - 1; t >= 0; t--) first[t] = std::min(first[t], first[t + 1]);
    std::vector<LibraryPart> parts(T);
    parallel([&](int t) { parseRange(first[t], first[t + 1], base + n, keepQual, parts[t]); });
    for (auto& part : parts)
        if (!part.ok) return false;
    // join in range order, each range copying its own arenas
    if (out.hoff.empty()) {
        out.hoff.push_back(0);
        out.soff.push_back(0);
    }
    std::vector<u64> h0(T + 1), s0(T + 1), q0(T + 1), r0(T + 1);
    h0[0] = out.headers.size();
    s0[0] = out.seqs.size();
    q0[0] = out.quals.size();
    r0[0] = out.size();
    for (int t = 0; t < T; t++) {
        h0[t + 1] = h0[t] + parts[t].headers.size();
        s0[t + 1] = s0[t] + parts[t].seqs.size();
        q0[t + 1] = q0[t] + parts[t].quals.size();
        r0[t + 1] = r0[t] + parts[t].hend.size();
    }
    out.headers.resizeUninit(h0[T]);
    out.seqs.resizeUninit(s0[T]);
    out.quals.resizeUninit(q0[T]);
    out.hoff.resize(r0[T] + 1);
    out.soff.resize(r0[T] + 1);
    parallel([&](int t) {
        LibraryPart& part = parts[t];
        if (!part.headers.empty()) std::memcpy(out.headers.data() + h0[t], part.headers.data(), part.headers.size());
        if (!part.seqs.empty()) std::memcpy(out.seqs.data() + s0[t], part.seqs.data(), part.seqs.size());
        if (!part.quals.empty()) std::memcpy(out.quals.data() + q0[t], part.quals.data(), part.quals.size());
        for (size_t i = 0; i < part.hend.size(); i++) {
            out.hoff[r0[t] + 1 + i] = h0[t] + part.hend[i];
            out.soff[r0[t] + 1 + i] = s0[t] + part.send[i];
        }
        part = LibraryPart();
    });
    out.fastq = true;
    return true;
}

// Mates of two libraries interleaved, up to the shorter one.
Library interleave(const Library& a, const Library& b) {
    Library lib;
    lib.hoff.push_back(0);
    lib.soff.push_back(0);
    lib.fileOf.push_back(0);
    const u64 n = std::min(a.size(), b.size());
    auto add = [&](const Library& x, u64 i) {
        lib.headers.append(x.headers.data() + x.hoff[i], x.hoff[i + 1] - x.hoff[i]);
        lib.hoff.push_back(lib.headers.size());
        lib.seqs.append(x.seqs.data() + x.soff[i], x.soff[i + 1] - x.soff[i]);
        if (!x.quals.empty()) lib.quals.append(x.quals.data() + x.soff[i], x.soff[i + 1] - x.soff[i]);
        lib.soff.push_back(lib.seqs.size());
    };
    for (u64 i = 0; i < n; i++) {
        add(a, i);
        add(b, i);
    }
    lib.fastq = a.fastq;
    return lib;
}

Library readLibrarySequential(const std::vector<std::string>& files, bool pairedEnd, bool keepQual) {
    Library lib;
    std::string h1, s1, q1, h2, s2, q2;
    if (pairedEnd && files.size() == 2) {
        // include/readlibraryio.hpp:130-190: mates of two files interleaved
        RecordReader a(files[0]), b(files[1]);
        lib.fileOf.push_back(0);
        while (a.next(h1, s1, q1)) {
            if (!b.next(h2, s2, q2)) break;
            lib.add(h1, s1, q1, keepQual);
            lib.add(h2, s2, q2, keepQual);
        }
        lib.fastq = a.fastq;
    } else {
        for (const auto& f : files) {
            RecordReader a(f);
            lib.fileOf.push_back(u32(lib.size()));
            while (a.next(h1, s1, q1)) lib.add(h1, s1, q1, keepQual);
            lib.fastq = a.fastq;
        }
    }
    if (lib.hoff.empty()) {
        lib.hoff.push_back(0);
        lib.soff.push_back(0);
    }
    return lib;
}

// All reads of the inputs in read-id order (include/readlibraryio.hpp:130-190):
// four-line FASTQ files in parallel, anything else sequentially.
Library readLibrary(const std::vector<std::string>& files, bool pairedEnd, bool keepQual, int threads) {
    if (threads > 1) {
        if (pairedEnd && files.size() == 2) {
            Library a, b;
            if (readFastq4(files[0], keepQual, threads, a) && readFastq4(files[1], keepQual, threads, b))
                return interleave(a, b);
        } else {
            Library lib;
            bool ok = true;
            for (const auto& f : files) {
                lib.fileOf.push_back(u32(lib.size()));
                if (!readFastq4(f, keepQual, threads, lib)) {
                    ok = false;
                    break;
                }
            }
            if (ok) {
                if (lib.hoff.empty()) {
                    lib.hoff.push_back(0);
                    lib.soff.push_back(0);
                }
                return lib;
            }
        }
    }
    return readLibrarySequential(files, pairedEnd, keepQual);
}

}  // namespace

// ---------------------------------------------------------------------------
// Packed read store (include/chunkedreadstorageconstruction.hpp:67-92,216-300)
// ---------------------------------------------------------------------------

namespace {

void encodeQualityRow(u32* out, const char* q, int len, int bits) {
    if (bits == 8) {
        std::memcpy(out, q, size_t(len));
        return;
    }
    // bins of every character: Q = char - 33, bin edges 13 (1 bit) and 4, 13, 20 (2 bits)
    static const std::array<std::array<u8, 256>, 2> kBins = [] {
        std::array<std::array<u8, 256>, 2> t{};
        for (int c = 0; c < 256; c++) {
            const int Q = int(static_cast<signed char>(c)) - 33;
            t[0][size_t(c)] = Q <= 13 ? 0 : 1;
            t[1][size_t(c)] = Q <= 4 ? 0 : Q <= 13 ? 1 : Q <= 20 ? 2 : 3;
        }
        return t;
    }();
    const u8* bins = kBins[size_t(bits - 1)].data();
    const int per = 32 / bits;
    const int words = (bits * len + 31) / 32;
    for (int w = 0; w < words; w++) {
        u32 data = 0;
        const int e = std::min(len, (w + 1) * per);
        for (int p = w * per; p < e; p++) data = (data << bits) | bins[static_cast<unsigned char>(q[p])];
        if (w == words - 1) data <<= ((w + 1) * per - len) * bits;
        out[w] = data;
    }
}

// Packs every read; non-ACGT characters become A,C,G,T in turn with a counter
// that restarts at every block of 65536 reads counted from each file start.
std::unique_ptr<care_readset> packLibrary(const Library& lib, int qualBits, int threads) {
    auto rs = std::make_unique<care_readset>();
    rs->n = lib.size();
    int maxLen = 0;
    for (u64 i = 0; i < rs->n; i++) maxLen = std::max(maxLen, lib.len(i));
    rs->maxLen = maxLen;
    rs->pitch = (maxLen + 15) / 16;
    rs->qbits = qualBits;
    rs->qpitch = qualBits ? (qualBits * maxLen + 31) / 32 : 0;
    rs->seqs.assign(rs->n * rs->pitch, 0);
    rs->lens.resize(rs->n);
    rs->ambig.assign(rs->n, 0);
    if (qualBits) rs->quals.assign(rs->n * u64(rs->qpitch), 0);
    // batches of the N rule
    std::vector<std::pair<u64, u64>> batches;
    std::vector<u64> starts(lib.fileOf.begin(), lib.fileOf.end());
    starts.push_back(rs->n);
    for (size_t f = 0; f + 1 < starts.size(); f++)
        for (u64 b = starts[f]; b < starts[f + 1]; b += 65536) batches.emplace_back(b, std::min(starts[f + 1], b + 65536));
    std::atomic<size_t> next{0};
    // 2-bit codes of the characters, either case; 4 = anything else
    static const std::array<u8, 256> kCode = [] {
        std::array<u8, 256> t{};
        t.fill(4);
        t['A'] = t['a'] = 0;
        t['C'] = t['c'] = 1;
        t['G'] = t['g'] = 2;
        t['T'] = t['t'] = 3;
        return t;
    }();
    auto work = [&]() {
        for (;;) {
            const size_t bi = next.fetch_add(1);
            if (bi >= batches.size()) break;
            int ncount = 0;
            for (u64 i = batches[bi].first; i < batches[bi].second; i++) {
                const int len = lib.len(i);
                const char* s = lib.seq(i);
                bool amb = false;
                u32* row = rs->seqs.data() + i * rs->pitch;
                for (int w0 = 0; w0 < len; w0 += 16) {
                    const int e = std::min(16, len - w0);
                    u32 word = 0;
                    for (int j = 0; j < e; j++) {
                        u32 c = kCode[static_cast<unsigned char>(s[w0 + j])];
                        if (c > 3) {
                            c = u32(ncount);
                            ncount = (ncount + 1) & 3;
                            amb = true;
                        }
                        word |= c << (30 - 2 * j);
                    }
                    row[w0 >> 4] = word;
                }
                rs->ambig[i] = amb;
                rs->lens[i] = len;
                if (qualBits) encodeQualityRow(rs->quals.data() + i * rs->qpitch, lib.quals.data() + lib.soff[i], len, qualBits);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); t++) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    return rs;
}

// ---------------------------------------------------------------------------
// Options (src/options.cpp:705-895; defaults include/options.hpp:36-98)
// ---------------------------------------------------------------------------
struct HostOptions {
    care_options eng{};
    std::vector<std::string> inputs, outputs;
    std::string outdir;
    int qualityBits = 8;
    int threads = 1;
    u64 fixedNumberOfReads = 0;
    bool labels = false;
    bool gzoutput = false;
    bool showProgress = false;
    std::vector<int> gpus;
    int batchsize = 0;
    bool autotune = true;  // time launch shapes on a sample of anchors before STEP 2
    int candResultsPerAnchor = 64;  // room for candidate corrections per anchor of a chunk
    float hashLoadFactor = 0.8f;    // --hashloadfactor: load of the lookups in a written hash-table file
    u64 memTotal = 0;               // -m/--memTotal in bytes (0: not given)
    std::string tempdir;            // --tempdir (default: the output directory)
    std::string forestAnchor, forestCands;
    // checkpoint files of the reference (preprocessed reads, hash tables)
    std::string saveReads, loadReads, saveTables, loadTables;
    std::string gateFile;       // learned gate parameters (correctionType 3)
    std::string gateFileCands;  // candidate gate parameters (correctionTypeCands 3; default: the anchor gate)
    uint32_t maxTreesAnchor = 0xFFFFFFFFu, maxTreesCands = 0xFFFFFFFFu;
    bool haveCoverage = false, havePairmode = false;
};

std::vector<int> parseIntList(const std::string& s) {
    std::vector<int> v;
    size_t p = 0;
    while (p < s.size()) {
        size_t e = s.find(',', p);
        if (e == std::string::npos) e = s.size();
        v.push_back(std::stoi(s.substr(p, e - p)));
        p = e + 1;
    }
    return v;
}

// byte count with an optional K, M or G suffix (src/options.cpp:239-265)
u64 parseMemory(const std::string& v) {
    if (v.empty()) return 0;
    u64 factor = 1;
    std::string digits = v;
    switch (v.back()) {
        case 'K': factor = u64(1) << 10; break;
        case 'M': factor = u64(1) << 20; break;
        case 'G': factor = u64(1) << 30; break;
        default: break;
    }
    if (factor != 1) digits.pop_back();
    return factor * std::stoull(digits);
}

HostOptions parseOptions(int argc, char** argv) {
    HostOptions o;
    care_options_default(&o.eng);
    auto boolValue = [](const std::string& v) { return !(v == "false" || v == "0" || v == "no"); };
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string inlineVal;
        bool hasInline = false;
        if (a.rfind("--", 0) == 0) {
            const size_t eq = a.find('=');
            if (eq != std::string::npos) {
                inlineVal = a.substr(eq + 1);
                a = a.substr(0, eq);
                hasInline = true;
            }
        }
        auto value = [&]() -> std::string {
            if (hasInline) return inlineVal;
            if (i + 1 >= argc) throw std::runtime_error("option " + a + " needs a value");
            return argv[++i];
        };
        auto flag = [&]() -> bool { return hasInline ? boolValue(inlineVal) : true; };
        if (a == "-i" || a == "--inputfiles") o.inputs.push_back(value());
        else if (a == "-o" || a == "--outputfilenames") o.outputs.push_back(value());
        else if (a == "-d" || a == "--outdir") o.outdir = value();
        else if (a == "-c" || a == "--coverage") { o.eng.coverage = std::stof(value()); o.haveCoverage = true; }
        else if (a == "--pairmode") {
            const std::string m = value();
            if (m == "PE" || m == "pe") o.eng.paired_end = 1;
            else if (m == "SE" || m == "se") o.eng.paired_end = 0;
            else throw std::runtime_error("invalid pairmode " + m);
            o.havePairmode = true;
        } else if (a == "-g" || a == "--gpu") o.gpus = parseIntList(value());
        else if (a == "-k" || a == "--kmerlength") o.eng.kmer_length = std::stoi(value());
        else if (a == "-h" || a == "--hashmaps") o.eng.num_hash_maps = std::stoi(value());
        else if (a == "-t" || a == "--threads") o.threads = std::stoi(value());
        else if (a == "-q" || a == "--useQualityScores") o.eng.use_quality = flag();
        else if (a == "--qualityScoreBits") o.qualityBits = std::stoi(value());
        else if (a == "--excludeAmbiguous") o.eng.exclude_ambiguous = flag();
        else if (a == "--maxmismatchratio") o.eng.max_mismatch_ratio = std::stof(value());
        else if (a == "--minalignmentoverlap") o.eng.min_overlap = std::stoi(value());
        else if (a == "--minalignmentoverlapratio") o.eng.min_overlap_ratio = std::stof(value());
        else if (a == "--errorfactortuning") o.eng.error_factor = std::stof(value());
        else if (a == "--coveragefactortuning") o.eng.coverage_factor = std::stof(value());
        else if (a == "--candidateCorrection") o.eng.correct_candidates = flag();
        else if (a == "--candidateCorrectionNewColumns") o.eng.new_columns_to_correct = std::stoi(value());
        else if (a == "--correctionQualityLabels") o.labels = flag();
        else if (a == "--correctionType") o.eng.correction_type = std::stoi(value());
        else if (a == "--correctionTypeCands") o.eng.correction_type_cands = std::stoi(value());
        else if (a == "--ml-forestfile") o.forestAnchor = value();
        else if (a == "--ml-gatefile") o.gateFile = value();
        else if (a == "--ml-cands-gatefile") o.gateFileCands = value();
        else if (a == "--candidateLimit") o.eng.candidate_limit = std::stoi(value());
        else if (a == "--refinements") o.eng.max_refinements = std::stoi(value());
        else if (a == "--ml-cands-forestfile") o.forestCands = value();
        else if (a == "--thresholdAnchor") o.eng.threshold_anchor = std::stof(value());
        else if (a == "--thresholdCands") o.eng.threshold_cands = std::stof(value());
        else if (a == "--pairedFilterThreshold") o.eng.paired_filter_threshold = std::stof(value());
        else if (a == "--fixedNumberOfReads") o.fixedNumberOfReads = std::stoull(value());
        else if (a == "--maxForestTreesAnchor") {
            const int v = std::stoi(value());
            if (v > 0) o.maxTreesAnchor = uint32_t(v);  // src/options.cpp:136-145
        } else if (a == "--maxForestTreesCands") {
            const int v = std::stoi(value());
            if (v > 0) o.maxTreesCands = uint32_t(v);
        }
        else if (a == "--batchsize") o.batchsize = std::stoi(value());
        else if (a == "--candidateResultsPerAnchor") {
            o.candResultsPerAnchor = std::stoi(value());
            if (o.candResultsPerAnchor < 1) throw std::runtime_error("--candidateResultsPerAnchor must be >= 1");
        }
        else if (a == "--autotune") o.autotune = flag();  // --autotune=false keeps the default shapes
        else if (a == "--gzoutput") o.gzoutput = flag();
        else if (a == "-p" || a == "--showProgress") o.showProgress = flag();
        else if (a == "--enforceHashmapCount") (void)flag();
        else if (a == "--save-preprocessedreads-to") o.saveReads = value();
        else if (a == "--load-preprocessedreads-from") o.loadReads = value();
        else if (a == "--save-hashtables-to") o.saveTables = value();
        else if (a == "--load-hashtables-from") o.loadTables = value();
        else if (a == "--hashloadfactor") {
            o.hashLoadFactor = std::stof(value());
            if (!(o.hashLoadFactor > 0.f && o.hashLoadFactor <= 1.f))
                throw std::runtime_error("--hashloadfactor must lie in (0, 1]");
        }
        else if (a == "-m" || a == "--memTotal") o.memTotal = parseMemory(value());
        else if (a == "--tempdir") o.tempdir = value();
        // the host memory for hash tables and the reference's table
        // implementation choices do not apply to this build's device tables
        else if (a == "--memHashtables" ||
                 a == "--warpcore" || a == "--replicateGpuData" || a == "--singlehash" ||
                 a == "--samplingRateAnchor" || a == "--samplingRateCands")
            (void)value();
        else throw std::runtime_error("unknown option " + a);
    }
    return o;
}

// Page-locked host array from the engine (care_host_buffer_alloc).
template <class T>
struct PinnedArray {
    T* p = nullptr;
    explicit PinnedArray(size_t n) {
        void* v = nullptr;
        if (care_host_buffer_alloc(n * sizeof(T), &v) != CARE_OK)
            throw std::runtime_error(std::string("engine: ") + care_last_error());
        p = static_cast<T*>(v);
    }
    ~PinnedArray() { care_host_buffer_free(p); }
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    T* data() { return p; }
    T& operator[](size_t i) { return p[i]; }
};

// ---------------------------------------------------------------------------
// Correction results (include/correctedsequence.hpp:22-146 semantics)
// ---------------------------------------------------------------------------
// Result bytes held in memory up to a limit and appended to a temporary
// file in the temporary directory beyond it, the file mapped back for
// STEP 3 (what the reference's SerializedObjectStorage over a
// FileBackedUVector does with -m/--memTotal and --tempdir,
// include/serializedobjectstorage.hpp:15-80).  The file is unlinked as soon
// as it is created, so it disappears with the process.
class SpillBytes {
public:
    SpillBytes() = default;
    SpillBytes(const SpillBytes&) = delete;
    SpillBytes& operator=(const SpillBytes&) = delete;
    SpillBytes(SpillBytes&& o) noexcept { *this = std::move(o); }
    SpillBytes& operator=(SpillBytes&& o) noexcept {
        if (this != &o) {
            release();
            mem_ = std::move(o.mem_);
            limit_ = o.limit_;
            dir_ = std::move(o.dir_);
            fd_ = o.fd_;
            fileBytes_ = o.fileBytes_;
            map_ = o.map_;
            o.fd_ = -1;
            o.fileBytes_ = 0;
            o.map_ = nullptr;
        }
        return *this;
    }
    ~SpillBytes() { release(); }
    void configure(u64 limit, const std::string& dir) {
        limit_ = limit;
        dir_ = dir;
    }
    u64 size() const { return fileBytes_ + mem_.size(); }
    u64 spilledBytes() const { return fileBytes_; }
    void push_back(char c) { mem_.push_back(c); }
    void append(const char* p, size_t n) { mem_.insert(mem_.end(), p, p + n); }
    // back to n bytes; n never reaches into the spilled part (records are
    // settled whole)
    void truncate(u64 n) { mem_.resize(size_t(n - fileBytes_)); }
    // between records: beyond the limit the bytes in memory go to the file
    void settle() {
        if (mem_.size() > limit_) spill();
    }
    // all bytes, contiguous: memory, or the mapped file
    const char* finish() {
        if (fd_ < 0) return mem_.data();
        if (!map_) {
            spill();
            void* m = mmap(nullptr, size_t(fileBytes_), PROT_READ, MAP_PRIVATE, fd_, 0);
            if (m == MAP_FAILED) throw std::runtime_error("cannot map the spilled results");
            map_ = static_cast<const char*>(m);
        }
        return map_;
    }

private:
    void spill() {
        if (fd_ < 0) {
            std::string path = dir_ + "/care_results_XXXXXX";
            fd_ = mkstemp(&path[0]);
            if (fd_ < 0) throw std::runtime_error("cannot create a temporary file in " + dir_);
            unlink(path.c_str());
        }
        size_t done = 0;
        while (done < mem_.size()) {
            const ssize_t w = write(fd_, mem_.data() + done, mem_.size() - done);
            if (w <= 0) throw std::runtime_error("write error on a temporary file in " + dir_);
            done += size_t(w);
        }
        fileBytes_ += mem_.size();
        mem_.clear();
    }
    void release() {
        if (map_) munmap(const_cast<char*>(map_), size_t(fileBytes_));
        if (fd_ >= 0) close(fd_);
        map_ = nullptr;
        fd_ = -1;
    }
    std::vector<char> mem_;
    u64 limit_ = ~u64(0);
    std::string dir_ = ".";
    int fd_ = -1;
    u64 fileBytes_ = 0;
    const char* map_ = nullptr;
};

struct ResultStore {
    // one entry per stored result
    std::vector<u32> readId;
    std::vector<u8> flags;      // bit0 anchor, bit1 hq, bit2 useEdits
    std::vector<u64> dataOff;   // offset into data (edits: pos u16 + base; or full sequence)
    std::vector<u32> dataLen;   // number of edits or sequence length
    SpillBytes data;

    void add(u32 id, bool anchor, bool hq, const char* corrected, const char* original, int len, bool ambiguous) {
        // include/corrector.hpp:1645-1668 and 1717-1757
        u8 f = u8((anchor ? 1 : 0) | (hq ? 2 : 0));
        const u64 off = data.size();
        u32 cnt = 0;
        bool useEdits = false;
        if (!ambiguous) {
            const int maxEdits = len / 7;
            int edits = 0;
            for (int i = 0; i < len && edits <= maxEdits; i++) {
                if (corrected[i] != original[i]) {
                    data.push_back(char(i & 0xFF));
                    data.push_back(char((i >> 8) & 0xFF));
                    data.push_back(corrected[i]);
                    edits++;
                }
            }
            useEdits = edits <= maxEdits;
            if (useEdits) cnt = u32(edits);
            else data.truncate(off);
        }
        if (useEdits) {
            // HQ anchors without edits are not stored (src/correct_cpu.cpp:262-266)
            if (anchor && hq && cnt == 0) return;
            f |= 4;
        } else {
            data.append(corrected, size_t(len));
            cnt = u32(len);
        }
        readId.push_back(id);
        flags.push_back(f);
        dataOff.push_back(off);
        dataLen.push_back(cnt);
        data.settle();
    }
};

// Entries of several stores, each with a pointer to its bytes.
struct ResultView {
    std::vector<u32> readId;
    std::vector<u8> flags;
    std::vector<u32> dataLen;
    std::vector<const char*> ptr;
    // the entries of every store in store order, one thread per store
    void collect(std::vector<ResultStore>& parts) {
        std::vector<size_t> at(parts.size() + 1, readId.size());
        for (size_t i = 0; i < parts.size(); i++) at[i + 1] = at[i] + parts[i].readId.size();
        readId.resize(at.back());
        flags.resize(at.back());
        dataLen.resize(at.back());
        ptr.resize(at.back());
        std::vector<const char*> bases(parts.size());
        for (size_t i = 0; i < parts.size(); i++) bases[i] = parts[i].data.finish();  // may map a file
        std::vector<std::thread> pool;
        for (size_t i = 0; i < parts.size(); i++) {
            pool.emplace_back([&, i]() {
                ResultStore& s = parts[i];
                const size_t o = at[i];
                std::copy(s.readId.begin(), s.readId.end(), readId.begin() + long(o));
                std::copy(s.flags.begin(), s.flags.end(), flags.begin() + long(o));
                std::copy(s.dataLen.begin(), s.dataLen.end(), dataLen.begin() + long(o));
                for (size_t e = 0; e < s.dataOff.size(); e++) ptr[o + e] = bases[i] + s.dataOff[e];
                std::vector<u32>().swap(s.readId);
                std::vector<u8>().swap(s.flags);
                std::vector<u32>().swap(s.dataLen);
                std::vector<u64>().swap(s.dataOff);
            });
        }
        for (auto& t : pool) t.join();
    }
};

// src/correctionresultoutput.cpp:730-1023 (combineMultipleCorrectionResults1_rawtcs2);
// returns label 0 uncorrected, 1 lq anchor only, 2 lq with candidates, 3 hq
int combineRead(const ResultView& rsx, const std::vector<u64>& entries, std::string& seq) {
    if (entries.empty()) return 0;
    long anchorIdx = -1;
    for (u64 e : entries)
        if (rsx.flags[e] & 1) {
            anchorIdx = long(e);
            break;
        }
    if (anchorIdx < 0) return 0;
    auto apply = [&](u64 e, std::string& s) {
        if (rsx.flags[e] & 4) {
            const char* d = rsx.ptr[e];
            for (u32 k = 0; k < rsx.dataLen[e]; k++) {
                const int pos = int((unsigned char)d[3 * k]) | (int((unsigned char)d[3 * k + 1]) << 8);
                s[size_t(pos)] = d[3 * k + 2];
            }
        } else {
            s.assign(rsx.ptr[e], rsx.dataLen[e]);
        }
    };
    const u64 a = u64(anchorIdx);
    if (rsx.flags[a] & 2) {
        apply(a, seq);
        return 3;
    }
    if (entries.size() >= 3) {
        const u64 f0 = entries[0];
        std::string first = seq;
        apply(f0, first);
        for (size_t i = 1; i < entries.size(); i++) {
            const u64 e = entries[i];
            bool eq;
            const bool e0 = rsx.flags[f0] & 4, e1 = rsx.flags[e] & 4;
            if (e0 && e1) {
                eq = rsx.dataLen[f0] == rsx.dataLen[e] &&
                     std::memcmp(rsx.ptr[f0], rsx.ptr[e], 3 * size_t(rsx.dataLen[e])) == 0;
            } else if (!e0 && !e1) {
                eq = rsx.dataLen[f0] == rsx.dataLen[e] &&
                     std::memcmp(rsx.ptr[f0], rsx.ptr[e], rsx.dataLen[e]) == 0;
            } else {
                std::string other = seq;
                apply(e, other);
                eq = other == first;
            }
            if (!eq) return 0;
        }
        seq.swap(first);
        return 2;
    }
    apply(a, seq);
    return 1;
}

struct OutWriter {
    FILE* f = nullptr;
    gzFile gz = nullptr;
    std::string buf;
    OutWriter(const std::string& path, bool gzip) {
        if (gzip) {
            gz = gzopen(path.c_str(), "wb");
            if (!gz) throw std::runtime_error("cannot open output file " + path);
        } else {
            f = std::fopen(path.c_str(), "wb");
            if (!f) throw std::runtime_error("cannot open output file " + path);
        }
    }
    void flush() {
        if (buf.empty()) return;
        if (f) std::fwrite(buf.data(), 1, buf.size(), f);
        else gzwrite(gz, buf.data(), unsigned(buf.size()));
        buf.clear();
    }
    ~OutWriter() {
        flush();
        if (f) std::fclose(f);
        if (gz) gzclose(gz);
    }
};

int runMain(int argc, char** argv) {
    HostOptions o = parseOptions(argc, argv);
    if (o.inputs.empty() || o.outputs.empty() || o.outdir.empty() || !o.haveCoverage || !o.havePairmode) {
        std::fprintf(stderr, "Mandatory options: -i/--inputfiles, -o/--outputfilenames, -d/--outdir, -c/--coverage, --pairmode, -g/--gpu\n");
        return 0;  // the reference exits with status 0 on invalid options
    }
    if (o.gpus.empty()) {
        std::fprintf(stderr, "No GPU specified (-g). This build runs on one device, e.g. -g 0\n");
        return 0;
    }
    if (o.gpus.size() > 1) std::fprintf(stderr, "Using only GPU %d of the given list (single-device build)\n", o.gpus[0]);
    if (o.outputs.size() != 1 && o.outputs.size() != o.inputs.size())
        throw std::runtime_error("number of output files must be 1 or equal to the number of input files");
    if (o.eng.paired_end && o.inputs.size() > 2) throw std::runtime_error("paired-end input takes one or two files");
    if (o.eng.use_quality && o.qualityBits != 1 && o.qualityBits != 2 && o.qualityBits != 8)
        throw std::runtime_error("qualityScoreBits must be 1, 2 or 8");
    // output and temporary directories and their parents (std::filesystem::create_directories)
    if (o.tempdir.empty()) o.tempdir = o.outdir;
    for (const std::string* dir : {&o.outdir, &o.tempdir})
        for (size_t i = 1; i <= dir->size(); i++)
            if (i == dir->size() || (*dir)[i] == '/') mkdir(dir->substr(0, i).c_str(), 0755);
    const int threads = std::max(1, o.threads);

    // STEP 1
    std::printf("STEP 1: Database construction\n");
    const double t0 = now_s();
    Library lib = readLibrary(o.inputs, o.eng.paired_end != 0, true, threads);
    if (!lib.fastq) o.eng.use_quality = 0;
    const double tParsed = now_s();
    std::unique_ptr<care_readset> rs;
    if (!o.loadReads.empty()) {
        // the inputs are still read: the output is built from their records
        rs.reset(care_readset_load_preprocessed(o.loadReads.c_str(), o.eng.use_quality, o.qualityBits));
        if (!rs) throw std::runtime_error(care_host_last_error());
        if (rs->n != lib.size())
            throw std::runtime_error("preprocessed reads file holds " + std::to_string(rs->n) + " reads, the inputs " +
                                     std::to_string(lib.size()));
        std::printf("Loaded readstorage from file %s\n", o.loadReads.c_str());
    } else {
        rs = packLibrary(lib, o.eng.use_quality ? o.qualityBits : 0, threads);
    }
    const double tPacked = now_s();
    if (!o.saveReads.empty()) {
        std::printf("Saving reads to file %s\n", o.saveReads.c_str());
        if (care_readset_save(rs.get(), o.saveReads.c_str(), o.qualityBits) != 0)
            throw std::runtime_error(care_host_last_error());
        std::printf("Saved reads\n");
    }
    std::printf("Number of reads: %llu, max length %d (parsed in %.3fs, packed in %.3fs)\n", (unsigned long long)rs->n,
                rs->maxLen, tParsed - t0, tPacked - tParsed);
    care_context* ctx = nullptr;
    auto check = [&](int rc) {
        if (rc != CARE_OK) throw std::runtime_error(std::string("engine: ") + care_last_error());
    };
    check(care_context_create(o.gpus[0], &o.eng, &ctx));
    std::unique_ptr<care_context, int (*)(care_context*)> guard(ctx, care_context_destroy);
    if (o.batchsize > 0) check(care_set_launch_config(ctx, (o.batchsize + 1) & ~1, 0));
    if (o.eng.correction_type == 1) check(care_forest_load(ctx, 0, o.forestAnchor.c_str(), o.maxTreesAnchor));
    // gate parameters: CARE_MLP_PARAMS little-endian float32 values
    auto readGate = [](const std::string& path, const char* flag) {
        std::vector<float> params(CARE_MLP_PARAMS);
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open gate file '" + path + "' (" + flag + ")");
        const size_t got = std::fread(params.data(), sizeof(float), params.size(), f);
        const bool extra = std::fgetc(f) != EOF;
        std::fclose(f);
        if (got != params.size() || extra)
            throw std::runtime_error("gate file '" + path + "' must hold exactly CARE_MLP_PARAMS float32 values");
        return params;
    };
    if (o.eng.correction_type == 3) check(care_mlp_load(ctx, readGate(o.gateFile, "--ml-gatefile").data()));
    if (o.eng.correction_type_cands == 3 && !o.gateFileCands.empty())
        check(care_mlp_load_candidates(ctx, readGate(o.gateFileCands, "--ml-cands-gatefile").data()));
    if (o.eng.correct_candidates && o.eng.correction_type_cands == 1)
        check(care_forest_load(ctx, 1, o.forestCands.c_str(), o.maxTreesCands));
    check(care_readstore_upload(ctx, rs->n, rs->maxLen, rs->seqs.data(), rs->pitch, rs->lens.data(), rs->ambig.data(),
                                o.eng.use_quality ? rs->quals.data() : nullptr, rs->qpitch, rs->qbits));
    const double tUploaded = now_s();
    if (!o.loadTables.empty()) {
        int loaded = 0;
        if (care_minhash_load_file(ctx, o.loadTables.c_str(), o.eng.num_hash_maps, &loaded) != 0)
            throw std::runtime_error(care_host_last_error());
        int ratio = 0;
        check(care_minhash_geometry(ctx, &o.eng.kmer_length, &o.eng.num_hash_maps, &ratio));
        std::printf("Loaded %d hash tables from %s\n", loaded, o.loadTables.c_str());
    } else {
        check(care_minhash_build(ctx));
    }
    if (!o.saveTables.empty()) {
        std::printf("Saving minhasher to file %s\n", o.saveTables.c_str());
        if (care_minhash_save_file(ctx, o.saveTables.c_str(), 0, o.hashLoadFactor) != 0)
            throw std::runtime_error(care_host_last_error());
        std::printf("Saved minhasher\n");
    }
    const double tBuilt = now_s();
    std::printf("Device: context and read store upload %.3fs, minhash tables %.3fs\n", tUploaded - tPacked,
                tBuilt - tUploaded);
    std::printf("# elapsed time (STEP1): %.6fs\n", tBuilt - t0);

    // launch shapes timed on the first anchors (set-up, like the table
    // build; results do not depend on them)
    const u64 numAnchors = o.fixedNumberOfReads ? std::min<u64>(o.fixedNumberOfReads, rs->n) : rs->n;
    if (o.autotune && numAnchors >= 4096) {
        const double ta = now_s();
        care_tuning t{};
        check(care_autotune(ctx, 0, int(std::min<u64>(numAnchors, 65536) & ~u64(1)), &t));
        std::printf("Launch shapes: query %d hits in LDS, correction %d threads x %d per CU, MSA chunk %d "
                    "(tuned in %.3fs)\n",
                    t.query_lds_hits, t.correct_threads, t.correct_blocks_per_cu, t.msa_chunk, now_s() - ta);
    }
    // Result buffers in page-locked memory (the engine copies into them by
    // DMA), two sets: the device fills one while host threads turn the
    // other's corrections into edit lists.
    const int pitch = (rs->maxLen + 15) / 16 * 16;
    // chunks of 2^17 anchors (larger ones measured slower: each call's
    // result copy cannot overlap the next call), the last one split so that
    // the host's work on the final results is short
    const int chunk = int(std::min<u64>(u64(1) << 17, std::max<u64>(numAnchors + (numAnchors & 1), 2)));
    const int candCap = o.eng.correct_candidates ? chunk * o.candResultsPerAnchor : 0;
    struct Slot {
        PinnedArray<u8> status;
        PinnedArray<char> seqs;
        PinnedArray<u32> candIds;
        PinnedArray<char> candSeqs;
        u64 first = 0;
        int n = 0;
        int numCands = 0;
        Slot(size_t chunkN, size_t capN, size_t pitchN)
            : status(chunkN), seqs(chunkN * pitchN), candIds(std::max<size_t>(1, capN)),
              candSeqs(std::max<size_t>(1, capN) * pitchN) {}
    };
    std::vector<std::unique_ptr<Slot>> slots;
    for (int i = 0; i < 2; i++) slots.push_back(std::make_unique<Slot>(size_t(chunk), size_t(candCap), size_t(pitch)));

    // STEP 2
    std::printf("STEP 2: Error correction\n");
    const double t1 = now_s();
    std::vector<u8> hqAnchor(rs->n, 0);
    std::vector<ResultStore> anchorParts(static_cast<size_t>(threads)), candParts(static_cast<size_t>(threads));
    {
        // -m: what the library and the packed reads leave of the host memory
        // limit holds the results; beyond it they go to --tempdir
        const u64 held = lib.headers.size() + lib.seqs.size() + lib.quals.size() +
                         8 * (lib.hoff.size() + lib.soff.size()) + 4 * (rs->seqs.size() + rs->quals.size()) +
                         5 * rs->n;
        const u64 perStore = o.memTotal ? (o.memTotal > held ? (o.memTotal - held) / u64(2 * threads) : 0) : ~u64(0);
        for (auto& p : anchorParts) p.data.configure(perStore, o.tempdir);
        for (auto& p : candParts) p.data.configure(perStore, o.tempdir);
    }
    auto decode = [&](u64 id, std::string& out) {
        const int len = rs->lens[id];
        out.resize(size_t(len));
        const u32* row = rs->seqs.data() + id * rs->pitch;
        for (int p = 0; p < len; p++) out[size_t(p)] = "ACGT"[(row[p >> 4] >> (30 - 2 * (p & 15))) & 3u];
    };
    // edit lists of one filled slot, anchors and candidate results split
    // over the host threads (each thread appends to its own stores)
    auto processSlot = [&](Slot& sl) {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++) {
            pool.emplace_back([&, t]() {
                std::string original;
                const int a0 = int(u64(sl.n) * u64(t) / u64(threads)), a1 = int(u64(sl.n) * u64(t + 1) / u64(threads));
                for (int i = a0; i < a1; i++) {
                    if (!(sl.status[size_t(i)] & 1)) continue;
                    const u64 id = sl.first + u64(i);
                    const bool hq = (sl.status[size_t(i)] & 2) != 0;
                    if (hq) hqAnchor[id] = 1;
                    decode(id, original);
                    anchorParts[size_t(t)].add(u32(id), true, hq, sl.seqs.data() + size_t(i) * pitch, original.data(),
                                               rs->lens[id], rs->ambig[id] != 0);
                }
                const int c0 = int(u64(sl.numCands) * u64(t) / u64(threads));
                const int c1 = int(u64(sl.numCands) * u64(t + 1) / u64(threads));
                for (int i = c0; i < c1; i++) {
                    const u32 id = sl.candIds[size_t(i)];
                    decode(id, original);
                    candParts[size_t(t)].add(id, false, false, sl.candSeqs.data() + size_t(i) * pitch, original.data(),
                                             rs->lens[id], rs->ambig[id] != 0);
                }
            });
        }
        for (auto& th : pool) th.join();
    };
    // The device thread corrects chunk after chunk into the two slots; the
    // host side processes each slot as soon as it is filled.  A chunk whose
    // candidate results exceed the slot's capacity is corrected again in
    // halves.
    std::mutex mu;
    std::condition_variable cv;
    double deviceBusy = 0, hostBusy = 0, firstResult = -1;  // seconds, for the STEP 2 breakdown
    int deviceCalls = 0;
    int filled[2] = {0, 0};  // 1: results waiting for the host
    bool deviceDone = false;
    std::string deviceError;
    std::vector<int> order;  // slot of every filled batch, in order
    std::thread device([&]() {
        // pending anchor ranges as a stack, the next range on top
        std::vector<std::pair<u64, int>> work;
        const u64 tail = std::max<u64>(u64(chunk) / 8, 2);
        for (u64 first = 0; first < numAnchors;) {
            const u64 rem = numAnchors - first;
            u64 take = std::min<u64>(u64(chunk), rem);
            if (rem > tail && rem <= u64(chunk) + tail) take = rem - tail;  // a short final chunk
            int n = int(take);
            if (o.eng.paired_end && (n & 1)) n++;  // whole pairs
            if (first + u64(n) > rs->n) n = int(rs->n - first);
            work.emplace_back(first, n);
            first += u64(n);
        }
        std::reverse(work.begin(), work.end());
        int next = 0;
        while (!work.empty()) {
            const auto [first, n] = work.back();
            work.pop_back();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !filled[next]; });
            }
            Slot& sl = *slots[size_t(next)];
            care_batch_result r{};
            r.anchor_status = sl.status.data();
            r.anchor_sequences = sl.seqs.data();
            r.sequence_pitch = pitch;
            r.candidate_read_ids = sl.candIds.data();
            r.candidate_sequences = sl.candSeqs.data();
            r.candidate_capacity = candCap;
            const double tc = now_s();
            const int rc = care_correct_batch(ctx, first, n, &r);
            deviceBusy += now_s() - tc;
            deviceCalls++;
            if (firstResult < 0) firstResult = now_s() - t1;
            if (rc == CARE_ERR_CAPACITY && n > 2) {
                // candidate results beyond the slot: the two halves separately
                std::printf("Candidate corrections of reads %llu..%llu exceed the result buffer: corrected in halves\n",
                            (unsigned long long)first, (unsigned long long)(first + u64(n) - 1));
                int h = n / 2;
                if (o.eng.paired_end) h &= ~1;
                work.emplace_back(first + u64(h), n - h);
                work.emplace_back(first, h);
                continue;
            }
            std::lock_guard<std::mutex> lk(mu);
            if (rc != CARE_OK) {
                deviceError = care_last_error();
                break;
            }
            sl.first = first;
            sl.n = n;
            sl.numCands = r.num_candidates;
            filled[next] = 1;
            order.push_back(next);
            cv.notify_all();
            next ^= 1;
        }
        std::lock_guard<std::mutex> lk(mu);
        deviceDone = true;
        cv.notify_all();
    });
    for (size_t k = 0;; k++) {
        int slot;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return order.size() > k || deviceDone; });
            if (order.size() <= k) break;
            slot = order[k];
        }
        const double th = now_s();
        processSlot(*slots[size_t(slot)]);
        hostBusy += now_s() - th;
        std::lock_guard<std::mutex> lk(mu);
        filled[slot] = 0;
        cv.notify_all();
    }
    device.join();
    if (!deviceError.empty()) throw std::runtime_error("engine: " + deviceError);
    const double tLoop = now_s();
    ResultView anchors, cands;
    anchors.collect(anchorParts);
    cands.collect(candParts);
    u64 spilled = 0;
    for (auto& p : anchorParts) spilled += p.data.spilledBytes();
    for (auto& p : candParts) spilled += p.data.spilledBytes();
    if (spilled) std::printf("Correction results: %llu bytes kept in temporary files in %s\n",
                             (unsigned long long)spilled, o.tempdir.c_str());
    const double t2 = now_s();
    std::printf("Correction: %d device calls busy %.3fs (first results after %.3fs), host edit lists %.3fs, "
                "result collection %.3fs\n",
                deviceCalls, deviceBusy, firstResult, hostBusy, t2 - tLoop);
    std::printf("# elapsed time (STEP2): %.6fs\n", t2 - t1);
    std::printf("Correction throughput : ~%.1f reads/second.\n", double(numAnchors) / (t2 - t1));

    // STEP 3
    std::printf("STEP 3: Constructing output file(s)\n");
    // per-read entry lists; candidate results of reads corrected as HQ
    // anchors are dropped (include/corrector.hpp:1693-1697)
    std::vector<u32> count(rs->n + 1, 0);
    for (size_t e = 0; e < anchors.readId.size(); e++) count[anchors.readId[e]]++;
    for (size_t e = 0; e < cands.readId.size(); e++)
        if (!hqAnchor[cands.readId[e]]) count[cands.readId[e]]++;
    std::vector<u64> start(rs->n + 1, 0);
    for (u64 i = 0; i < rs->n; i++) start[i + 1] = start[i] + count[i];
    // merged store: anchors first, then candidates (offsets into one index space)
    const u64 na = anchors.readId.size();
    std::vector<u64> entry(start[rs->n]);
    std::vector<u64> fill(start.begin(), start.end() - 1);
    for (u64 e = 0; e < na; e++) entry[fill[anchors.readId[e]]++] = e;
    for (u64 e = 0; e < cands.readId.size(); e++)
        if (!hqAnchor[cands.readId[e]]) entry[fill[cands.readId[e]]++] = na + e;
    ResultView all = std::move(anchors);
    all.readId.insert(all.readId.end(), cands.readId.begin(), cands.readId.end());
    all.flags.insert(all.flags.end(), cands.flags.begin(), cands.flags.end());
    all.dataLen.insert(all.dataLen.end(), cands.dataLen.begin(), cands.dataLen.end());
    all.ptr.insert(all.ptr.end(), cands.ptr.begin(), cands.ptr.end());
    cands = ResultView();
    long long lastWithResult = -1;
    for (u64 i = 0; i < rs->n; i++)
        if (count[i]) lastWithResult = (long long)i;

    const bool perFile = o.outputs.size() > 1;
    std::vector<std::unique_ptr<OutWriter>> writers;
    for (const auto& name : o.outputs) writers.push_back(std::make_unique<OutWriter>(o.outdir + "/" + name, o.gzoutput));
    std::vector<u64> fileEnd;
    for (size_t f = 1; f < lib.fileOf.size(); f++) fileEnd.push_back(lib.fileOf[f]);
    fileEnd.push_back(rs->n);
    size_t fileIdx = 0;
    std::string seq;
    std::vector<u64> entries;
    for (u64 i = 0; i < rs->n; i++) {
        while (perFile && fileIdx + 1 < fileEnd.size() && i >= fileEnd[fileIdx]) fileIdx++;
        OutWriter& w = *writers[perFile ? fileIdx : 0];
        seq.assign(lib.seq(i), size_t(lib.len(i)));
        entries.assign(entry.begin() + long(start[i]), entry.begin() + long(start[i + 1]));
        const int label = combineRead(all, entries, seq);
        w.buf += lib.fastq ? '@' : '>';
        w.buf.append(lib.headers.data() + lib.hoff[i], lib.hoff[i + 1] - lib.hoff[i]);
        if (o.labels && (long long)i <= lastWithResult) {
            static const char* lab[4] = {" care:q=0", " care:q=1", " care:q=2", " care:q=3"};
            w.buf += lab[label];
        }
        w.buf += '\n';
        w.buf += seq;
        w.buf += '\n';
        if (lib.fastq) {
            w.buf += "+\n";
            w.buf.append(lib.quals.data() + lib.soff[i], size_t(lib.len(i)));
            w.buf += '\n';
        }
        if (w.buf.size() > (size_t(1) << 22)) w.flush();
    }
    writers.clear();
    const double t3 = now_s();
    std::printf("# elapsed time (STEP3): %.6fs\n", t3 - t2);
    std::printf("Construction of output file(s) finished.\n");
    return 0;
}

}  // namespace

extern "C" {

int care_host_main(int argc, char** argv) {
    try {
        return runMain(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "care-gpu: %s\n", e.what());
        return 1;
    }
}

const char* care_host_last_error(void) { return g_hostError.c_str(); }

care_readset* care_readset_load(const char* const* files, int num_files, int paired_end, int qual_bits) {
    try {
        if (!files || num_files <= 0) throw std::runtime_error("no input files");
        if (qual_bits != 0 && qual_bits != 1 && qual_bits != 2 && qual_bits != 8)
            throw std::runtime_error("qual_bits must be 0, 1, 2 or 8");
        std::vector<std::string> f(files, files + num_files);
        const unsigned hcp = std::thread::hardware_concurrency();
        Library lib = readLibrary(f, paired_end != 0, qual_bits != 0, int(hcp ? std::min(hcp, 16u) : 4u));
        if (!lib.fastq) qual_bits = 0;
        const unsigned hc = std::thread::hardware_concurrency();
        return packLibrary(lib, qual_bits, int(hc ? std::min(hc, 16u) : 4u)).release();
    } catch (const std::exception& e) {
        g_hostError = e.what();
        return nullptr;
    }
}

void care_readset_free(care_readset* rs) { delete rs; }
uint64_t care_readset_size(const care_readset* rs) { return rs ? rs->n : 0; }
int care_readset_max_length(const care_readset* rs) { return rs ? rs->maxLen : 0; }
int care_readset_seq_pitch(const care_readset* rs) { return rs ? rs->pitch : 0; }
int care_readset_qual_pitch(const care_readset* rs) { return rs ? rs->qpitch : 0; }
const uint32_t* care_readset_seqs(const care_readset* rs) { return rs ? rs->seqs.data() : nullptr; }
const int32_t* care_readset_lengths(const care_readset* rs) { return rs ? rs->lens.data() : nullptr; }
const uint8_t* care_readset_ambiguous(const care_readset* rs) { return rs ? rs->ambig.data() : nullptr; }
const uint32_t* care_readset_quals(const care_readset* rs) { return rs && !rs->quals.empty() ? rs->quals.data() : nullptr; }

}  // extern "C"
