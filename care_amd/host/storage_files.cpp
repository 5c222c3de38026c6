// The reference's on-disk formats for a read store and for minhash tables,
// so that files written by the reference's care-gpu load here and the
// other way round:
//
//   preprocessed reads  ChunkedReadStorage::saveToFile / loadFromFile
//                       (include/chunkedreadstorage.hpp:158-400), lengths as
//                       a LengthStore (include/lengthstorage.hpp:164-204)
//   hash tables         writeToStream / loadFromStream of the minhasher
//                       (include/gpu/fakegpuminhasher.cuh:491-525,
//                       include/ordinaryminhasher.hpp:303-339), per map a
//                       CpuReadOnlyMultiValueHashTable
//                       (include/cpuhashtable.hpp:627-649) whose key lookup
//                       is an AoSCpuSingleValueHashTable (:221-248)
//
// Fields are packed one after the other in host byte order (x86-64: little
// endian), exactly as the reference's stream.write calls lay them down.

#include "../../include/care_amd.h"
#include "../../include/care_host.h"
#include "readset.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using u8 = std::uint8_t;
using u16 = std::uint16_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

extern __attribute__((visibility("hidden"))) thread_local std::string g_hostError;

namespace {

struct OutFile {
    FILE* f = nullptr;
    std::string path;
    explicit OutFile(const std::string& p) : path(p) {
        f = std::fopen(p.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open '" + p + "' for writing");
    }
    ~OutFile() {
        if (f) std::fclose(f);
    }
    void bytes(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("write error on '" + path + "'");
    }
    template <class T>
    void put(T v) {
        bytes(&v, sizeof(T));
    }
    long tell() { return std::ftell(f); }
    void seek(long pos) {
        if (std::fseek(f, pos, SEEK_SET) != 0) throw std::runtime_error("seek error on '" + path + "'");
    }
    void close() {
        if (std::fclose(f) != 0) {
            f = nullptr;
            throw std::runtime_error("write error on '" + path + "'");
        }
        f = nullptr;
    }
};

// Reads with a bound on every count: sizes come from the file, so each
// array is checked against the bytes that remain before it is allocated.
struct InFile {
    FILE* f = nullptr;
    std::string path;
    u64 size = 0;
    explicit InFile(const std::string& p) : path(p) {
        f = std::fopen(p.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open '" + p + "'");
        std::fseek(f, 0, SEEK_END);
        size = u64(std::ftell(f));
        std::fseek(f, 0, SEEK_SET);
    }
    ~InFile() {
        if (f) std::fclose(f);
    }
    u64 remaining() { return size - u64(std::ftell(f)); }
    void bytes(void* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) throw std::runtime_error("'" + path + "' ends early");
    }
    template <class T>
    T get() {
        T v;
        bytes(&v, sizeof(T));
        return v;
    }
    template <class T>
    void array(std::vector<T>& v, u64 count, const char* what) {
        if (count > remaining() / sizeof(T))
            throw std::runtime_error("'" + path + "': " + what + " extends beyond the end of the file");
        v.resize(count);
        bytes(v.data(), count * sizeof(T));
    }
    void skip(u64 n) {
        if (n > remaining()) throw std::runtime_error("'" + path + "' ends early");
        std::fseek(f, long(n), SEEK_CUR);
    }
};

// --- LengthStore -----------------------------------------------------------
// Lengths minus the minimum, `bits` bits each, in one bit string numbered
// from the most significant bit of word 0; a value's most significant bit
// comes first.  bits = bit width of (max - min), 0 when all lengths agree.

int lengthBits(int diff) {
    int b = 0;
    while (b < 31 && (u32(diff) >> b) != 0) b++;
    return b;
}

std::vector<u32> packLengths(const std::vector<int32_t>& lens, int minL, int bits) {
    std::vector<u32> data((u64(bits) * lens.size() + 31) / 32, 0);
    for (u64 i = 0; i < lens.size(); i++) {
        const u32 v = u32(lens[i] - minL);
        for (int j = 0; j < bits; j++) {
            const u64 bit = u64(bits) * i + u64(j);
            if ((v >> (bits - 1 - j)) & 1u) data[bit / 32] |= 1u << (31 - bit % 32);
        }
    }
    return data;
}

int unpackLength(const std::vector<u32>& data, u64 i, int minL, int bits) {
    u32 v = 0;
    for (int j = 0; j < bits; j++) {
        const u64 bit = u64(bits) * i + u64(j);
        v = (v << 1) | ((data[bit / 32] >> (31 - bit % 32)) & 1u);
    }
    return minL + int(v);
}

// --- preprocessed reads ----------------------------------------------------

void saveReads(const care_readset& rs, const std::string& path, int qualBitsField) {
    OutFile out(path);
    int minL = 0, maxL = 0;
    if (rs.n) {
        minL = *std::min_element(rs.lens.begin(), rs.lens.end());
        maxL = *std::max_element(rs.lens.begin(), rs.lens.end());
    }
    const bool qual = rs.qbits != 0 && !rs.quals.empty();
    out.put<u64>(rs.n);
    out.put<int32_t>(maxL);  // upper bound first, lower second (chunkedreadstorage.hpp:255-256)
    out.put<int32_t>(minL);
    out.put<u8>(qual ? 1 : 0);
    out.put<int32_t>(qual ? rs.qbits : qualBitsField);
    const long sizesAt = out.tell();
    for (int i = 0; i < 4; i++) out.put<u64>(0);

    // LengthStore<u32>
    const int bits = lengthBits(maxL - minL);
    const std::vector<u32> ldata = packLengths(rs.lens, minL, bits);
    out.put<int32_t>(32);
    out.put<int32_t>(bits);
    out.put<int32_t>(minL);
    out.put<int32_t>(maxL);
    out.put<u32>(bits ? u32((u64(1) << bits) - 1) : 0u);
    out.put<int64_t>(int64_t(rs.n));
    out.put<u64>(ldata.size());
    out.put<u64>(ldata.size() * sizeof(u32));
    out.bytes(ldata.data(), ldata.size() * sizeof(u32));
    const u64 lengthsBytes = 4 * sizeof(int32_t) + sizeof(u32) + sizeof(int64_t) + 2 * sizeof(u64) +
                             ldata.size() * sizeof(u32);

    out.put<u64>(u64(rs.pitch));
    out.put<u64>(rs.n * u64(rs.pitch));
    out.bytes(rs.seqs.data(), rs.n * u64(rs.pitch) * sizeof(u32));
    const u64 seqBytes = 2 * sizeof(u64) + rs.n * u64(rs.pitch) * sizeof(u32);

    const u64 qpitch = qual ? u64(rs.qpitch) : 0;
    out.put<u64>(qpitch);
    out.put<u64>(rs.n * qpitch);
    if (qual) out.bytes(rs.quals.data(), rs.n * qpitch * sizeof(u32));
    const u64 qualBytes = 2 * sizeof(u64) + rs.n * qpitch * sizeof(u32);

    std::vector<u32> ambigIds;
    for (u64 i = 0; i < rs.n; i++)
        if (rs.ambig[i]) ambigIds.push_back(u32(i));
    out.put<u64>(ambigIds.size());
    out.bytes(ambigIds.data(), ambigIds.size() * sizeof(u32));
    const u64 ambigBytes = sizeof(u64) + ambigIds.size() * sizeof(u32);

    out.seek(sizesAt);
    out.put<u64>(lengthsBytes);
    out.put<u64>(seqBytes);
    out.put<u64>(qualBytes);
    out.put<u64>(ambigBytes);
    out.close();
}

std::unique_ptr<care_readset> loadReads(const std::string& path, bool useQuality, int qualBits) {
    InFile in(path);
    auto rs = std::make_unique<care_readset>();
    const u64 n = in.get<u64>();
    (void)in.get<int32_t>();  // length bounds, repeated in the LengthStore
    (void)in.get<int32_t>();
    const bool hasQual = in.get<u8>() != 0;
    const int fileQualBits = in.get<int32_t>();
    (void)in.get<u64>();  // section sizes: lengths, sequences
    (void)in.get<u64>();
    const u64 qualBytes = in.get<u64>();
    (void)in.get<u64>();  // ambiguous ids

    if (n > 0xFFFFFFFFull) throw std::runtime_error("'" + path + "': more reads than 32-bit read ids hold");
    if (in.get<int32_t>() != 32) throw std::runtime_error("'" + path + "': length store is not 32-bit");
    const int bits = in.get<int32_t>();
    const int minL = in.get<int32_t>();
    const int maxL = in.get<int32_t>();
    (void)in.get<u32>();  // bit mask of `bits` ones
    const int64_t numLengths = in.get<int64_t>();
    const u64 rawElements = in.get<u64>();
    const u64 rawBytes = in.get<u64>();
    if (bits < 0 || bits > 31 || minL < 0 || maxL < minL || numLengths != int64_t(n) ||
        rawBytes != rawElements * sizeof(u32) || rawElements < (u64(bits) * n + 31) / 32)
        throw std::runtime_error("'" + path + "': malformed length store");
    std::vector<u32> ldata;
    in.array(ldata, rawElements, "length store");

    rs->n = n;
    rs->lens.resize(n);
    for (u64 i = 0; i < n; i++) {
        rs->lens[i] = bits ? unpackLength(ldata, i, minL, bits) : minL;
        if (rs->lens[i] > maxL) throw std::runtime_error("'" + path + "': read length above the stored maximum");
    }
    rs->maxLen = maxL;

    const u64 pitch = in.get<u64>();
    const u64 numSeqWords = in.get<u64>();
    if (pitch < u64(maxL + 15) / 16 || pitch > (1u << 20) || numSeqWords != n * pitch)
        throw std::runtime_error("'" + path + "': malformed sequence section");
    rs->pitch = int(pitch);
    in.array(rs->seqs, numSeqWords, "sequence section");

    if (useQuality && hasQual) {
        if (fileQualBits != qualBits)
            throw std::runtime_error("preprocessed reads file uses " + std::to_string(fileQualBits) +
                                     " bits per quality score, but setting requires " + std::to_string(qualBits) +
                                     " bits");
        const u64 qpitch = in.get<u64>();
        const u64 numQualWords = in.get<u64>();
        if (qpitch < (u64(qualBits) * u64(maxL) + 31) / 32 || qpitch > (1u << 22) || numQualWords != n * qpitch)
            throw std::runtime_error("'" + path + "': malformed quality section");
        rs->qbits = qualBits;
        rs->qpitch = int(qpitch);
        in.array(rs->quals, numQualWords, "quality section");
    } else if (useQuality && !hasQual) {
        throw std::runtime_error("Quality scores expected in preprocessed reads file to load, but none are present");
    } else {
        in.skip(qualBytes);
    }

    const u64 numAmbig = in.get<u64>();
    std::vector<u32> ambigIds;
    in.array(ambigIds, numAmbig, "ambiguous read ids");
    rs->ambig.assign(n, 0);
    for (u32 id : ambigIds) {
        if (id >= n) throw std::runtime_error("'" + path + "': ambiguous read id beyond the read count");
        rs->ambig[id] = 1;
    }
    return rs;
}

// --- hash tables ------------------------------------------------------------
// One lookup slot: std::pair<kmer, std::pair<read_number, BucketSize>>, 16
// bytes with 2 bytes of padding; empty slots hold all-ones fields.
struct Slot {
    u64 key;
    u32 offset;
    u16 count;
    u16 pad;
};
static_assert(sizeof(Slot) == 16, "lookup slot layout");

// hashers::MurmurHash<std::uint64_t>::hash, the MurmurHash3 64-bit finalizer
inline u64 fmix64(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// One map: its values in key order, then the lookup built as the reference
// builds it (include/cpuhashtable.hpp:63-67,96-121,516-545): capacity
// size / load computed in float, keys inserted in ascending order at
// murmur(key) % capacity with linear probing.
void writeMap(OutFile& out, const std::vector<u64>& keys, const std::vector<u32>& counts,
              const std::vector<u32>& values, float load) {
    out.put<u64>(values.size());
    out.bytes(values.data(), values.size() * sizeof(u32));
    const u64 size = keys.size();
    const u64 capacity = u64(float(size) / load);
    if (capacity < size) throw std::runtime_error("hash table capacity below its key count");
    std::vector<Slot> slots(capacity, Slot{~u64(0), ~u32(0), u16(0xFFFF), 0});
    u64 maxProbes = 0, offset = 0;
    for (u64 i = 0; i < size; i++) {
        u64 pos = fmix64(keys[i]) % capacity, probes = 0;
        while (slots[pos].key != ~u64(0)) {
            pos = pos + 1 == capacity ? 0 : pos + 1;
            probes++;
        }
        slots[pos] = Slot{keys[i], u32(offset), u16(counts[i]), 0};
        maxProbes = std::max(maxProbes, probes);
        offset += counts[i];
    }
    out.put<float>(load);
    out.put<u64>(size);  // numKeys
    out.put<u64>(maxProbes);
    out.put<u64>(size);
    out.put<u64>(capacity);
    out.put<u64>(capacity);  // elements of the storage vector
    out.bytes(slots.data(), slots.size() * sizeof(Slot));
}

struct LoadedMap {
    std::vector<u64> keys;
    std::vector<u32> offsets, counts, values;
};

void readMap(InFile& in, LoadedMap& m) {
    const u64 numValues = in.get<u64>();
    in.array(m.values, numValues, "table values");
    (void)in.get<float>();  // load factor
    (void)in.get<u64>();    // numKeys
    (void)in.get<u64>();    // maxProbes
    (void)in.get<u64>();    // size
    const u64 capacity = in.get<u64>();
    const u64 elements = in.get<u64>();
    if (elements != capacity) throw std::runtime_error("'" + in.path + "': malformed hash table lookup");
    std::vector<Slot> slots;
    in.array(slots, elements, "hash table lookup");
    std::vector<const Slot*> used;
    for (const Slot& s : slots)
        if (!(s.key == ~u64(0) && s.offset == ~u32(0) && s.count == 0xFFFF)) used.push_back(&s);
    std::sort(used.begin(), used.end(), [](const Slot* a, const Slot* b) { return a->key < b->key; });
    for (const Slot* s : used) {
        if (s->count == 0) continue;
        if (u64(s->offset) + s->count > numValues)
            throw std::runtime_error("'" + in.path + "': hash table entry beyond its values");
        m.keys.push_back(s->key);
        m.offsets.push_back(s->offset);
        m.counts.push_back(s->count);
        // each key's values ascending, as the device lists are merged
        std::sort(m.values.begin() + s->offset, m.values.begin() + s->offset + s->count);
    }
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_hostError = e.what();
        return 1;
    }
}

}  // namespace

extern "C" {

int care_readset_save(const care_readset* rs, const char* path, int qual_bits) {
    return guarded([&] {
        if (!rs || !path) throw std::runtime_error("null argument");
        saveReads(*rs, path, qual_bits);
    });
}

care_readset* care_readset_load_preprocessed(const char* path, int use_quality, int qual_bits) {
    try {
        if (!path) throw std::runtime_error("null path");
        if (use_quality && qual_bits != 1 && qual_bits != 2 && qual_bits != 8)
            throw std::runtime_error("qual_bits must be 1, 2 or 8");
        return loadReads(path, use_quality != 0, qual_bits).release();
    } catch (const std::exception& e) {
        g_hostError = e.what();
        return nullptr;
    }
}

int care_minhash_save_file(care_context* ctx, const char* path, int results_per_map, float load_factor) {
    return guarded([&] {
        if (!ctx || !path) throw std::runtime_error("null argument");
        if (!(load_factor > 0.f && load_factor <= 1.f)) throw std::runtime_error("load factor must be in (0, 1]");
        int k = 0, numMaps = 0, threshold = 0;
        if (care_minhash_geometry(ctx, &k, &numMaps, &threshold) != CARE_OK) throw std::runtime_error(care_last_error());
        OutFile out(path);
        out.put<int32_t>(k);
        out.put<int32_t>(results_per_map > 0 ? results_per_map : threshold);
        out.put<float>(load_factor);
        out.put<int32_t>(numMaps);
        for (int m = 0; m < numMaps; m++) {
            uint64_t nk = 0, nv = 0;
            if (care_minhash_map_size(ctx, m, &nk, &nv) != CARE_OK) throw std::runtime_error(care_last_error());
            std::vector<u64> keys(nk);
            std::vector<u32> offsets(nk), counts(nk), values(nv);
            if (care_minhash_export_map(ctx, m, keys.data(), offsets.data(), counts.data(), values.data()) != CARE_OK)
                throw std::runtime_error(care_last_error());
            for (u32 c : counts)
                if (c > 0xFFFF) throw std::runtime_error("bucket larger than the file format's 16-bit counts");
            writeMap(out, keys, counts, values, load_factor);
        }
        out.close();
    });
}

int care_minhash_load_file(care_context* ctx, const char* path, int max_maps, int* loaded_maps) {
    return guarded([&] {
        if (!ctx || !path) throw std::runtime_error("null argument");
        InFile in(path);
        const int k = in.get<int32_t>();
        const int threshold = in.get<int32_t>();
        (void)in.get<float>();  // load factor of the file's lookups
        const int numMaps = in.get<int32_t>();
        if (k < 1 || k > 31 || numMaps < 1 || threshold < 1)
            throw std::runtime_error("'" + std::string(path) + "': malformed hash table header");
        const int toLoad = std::min(numMaps, max_maps > 0 ? max_maps : numMaps);
        std::vector<LoadedMap> maps(size_t(std::min(toLoad, 64)));
        if (toLoad > 64) throw std::runtime_error("more than 64 hash tables");
        int maxCount = 1;
        for (auto& m : maps) {
            readMap(in, m);
            for (u32 c : m.counts) maxCount = std::max(maxCount, int(c));
        }
        std::vector<care_map_view> views(maps.size());
        for (size_t i = 0; i < maps.size(); i++)
            views[i] = care_map_view{maps[i].keys.size(), maps[i].keys.data(), maps[i].offsets.data(),
                                     maps[i].counts.data(), maps[i].values.size(), maps[i].values.data()};
        if (care_minhash_import(ctx, views.data(), toLoad, k, std::min(65535, std::max(threshold, maxCount))) !=
            CARE_OK)
            throw std::runtime_error(care_last_error());
        if (loaded_maps) *loaded_maps = toLoad;
    });
}

}  // extern "C"
