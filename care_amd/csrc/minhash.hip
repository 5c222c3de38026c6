// Minhash signatures, table construction and candidate queries on the device.
//
// Semantics follow the reference CPU minhasher:
//   signature: include/cpusequencehasher.hpp:88-125 with the canonical k-mer
//              walk of include/sequencehelpers.hpp:844-926 and the 64-bit
//              murmur finalizer of hpc_helpers hashers.cuh:129-137;
//   tables:    include/groupbykey.hpp:99-249 (keys with fewer than 2 or more
//              than cap values keep no values, include/config.hpp:32-34);
//   query:     include/ordinaryminhasher.hpp:135-221 followed by
//              include/corrector.hpp:571-650 (sort, unique, drop self, drop
//              ambiguous candidates under --excludeAmbiguous).
//
// The reference GPU path stores warpcore multi-value tables and compacts
// them (include/gpu/gpuhashtable.cuh:314-760).  Here every map is built with
// one counting pass into a scratch open-addressing table, an offset pass, a
// scatter pass and a per-bucket sort, then re-inserted into a compact table
// that only holds keys with 2..cap values.

#include "device_common.h"

namespace care_engine {

// ---------------------------------------------------------------------------
// Signatures: one block computes the h signature values of one read.
// Canonical k-mers go to LDS first; thread t then folds hash function
// t % h over k-mer positions t / h, t / h + blockDim / h, ...
// ---------------------------------------------------------------------------
__device__ void block_signature(const u32* enc, int len, int k, int h, u64* canon /*LDS, >= len*/,
                                u64* partial /*LDS, blockDim*/, u64* sig /*LDS, h*/) {
    const u64 kmask = ~u64(0) >> (2 * (32 - k));
    const int nk = len >= k ? len - k + 1 : 0;
    for (int p = threadIdx.x; p < nk; p += blockDim.x) {
        u64 fw = 0, rc = 0;
        for (int i = 0; i < k; i++) {
            const u64 b = base_at(enc, p + i);
            fw = (fw << 2) | b;
            rc = (rc >> 2) | ((3ull - b) << (2 * (k - 1)));
        }
        fw &= kmask;
        canon[p] = fw < rc ? fw : rc;
    }
    __syncthreads();
    const int parts = blockDim.x / h;
    for (int j0 = 0; j0 < h; j0 += blockDim.x) {
        const int t = threadIdx.x;
        u64 m = ~u64(0);
        int j, part, step;
        if (h <= int(blockDim.x)) {
            j = t % h;
            part = t / h;
            step = parts;
        } else {
            j = j0 + t;
            part = 0;
            step = 1;
        }
        const bool active = (h <= int(blockDim.x)) ? (part < parts) : (j < h);
        if (active) {
            for (int p = part; p < nk; p += step) {
                const u64 v = murmur64(canon[p] + u64(j));
                m = v < m ? v : m;
            }
        }
        partial[t] = m;
        __syncthreads();
        if (h <= int(blockDim.x)) {
            if (t < h) {
                u64 r = partial[t];
                for (int q = 1; q < parts; q++) {
                    const u64 v = partial[q * h + t];
                    r = v < r ? v : r;
                }
                sig[t] = r & kmask;
            }
        } else {
            if (j < h) sig[j] = partial[t] & kmask;
        }
        __syncthreads();
        if (h <= int(blockDim.x)) break;
    }
}

__global__ void k_signatures(DevReadStore rs, u64 first, int n, int k, int h, u64* out) {
    HIP_DYNAMIC_SHARED(u64, lds64)
    u64* canon = lds64;
    u64* partial = canon + rs.maxLen;
    u64* sig = partial + blockDim.x;
    for (u64 r = blockIdx.x; r < u64(n); r += gridDim.x) {
        const u64 id = first + r;
        block_signature(rs.seqs + id * rs.pitch, rs.lens[id], k, h, canon, partial, sig);
        for (int j = threadIdx.x; j < h; j += blockDim.x) out[r * h + j] = sig[j];
        __syncthreads();
    }
}

void launch_signatures(const DevReadStore& rs, u64 first, int n, int k, int h, u64* out, hipStream_t stream) {
    if (n <= 0) return;
    const int threads = 256;
    const int blocks = n < 65536 ? n : 65536;
    const size_t lds = sizeof(u64) * (size_t(rs.maxLen) + threads + h);
    hipLaunchKernelGGL(k_signatures, dim3(blocks), dim3(threads), lds, stream, rs, first, n, k, h, out);
}

// ---------------------------------------------------------------------------
// Table construction for one map
// ---------------------------------------------------------------------------
__global__ void k_fill_u64(u64* p, u64 n, u64 v) {
    for (u64 i = blockIdx.x * u64(blockDim.x) + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(u32* p, u64 n, u32 v) {
    for (u64 i = blockIdx.x * u64(blockDim.x) + threadIdx.x; i < n; i += u64(gridDim.x) * blockDim.x) p[i] = v;
}

__device__ __forceinline__ u64 find_or_insert(u64* keys, u64 mask, u64 key) {
    u64 slot = murmur64(key) & mask;
    for (u64 probe = 0; probe <= mask; probe++) {
        const u64 old = atomicCAS(reinterpret_cast<unsigned long long*>(&keys[slot]), (unsigned long long)kEmptyKey,
                                  (unsigned long long)key);
        if (old == kEmptyKey || old == key) return slot;
        slot = (slot + 1) & mask;
    }
    return ~u64(0);
}

__device__ __forceinline__ u64 find_slot(const u64* keys, u64 mask, u64 key) {
    u64 slot = murmur64(key) & mask;
    for (u64 probe = 0; probe <= mask; probe++) {
        const u64 cur = keys[slot];
        if (cur == key) return slot;
        if (cur == kEmptyKey) return ~u64(0);
        slot = (slot + 1) & mask;
    }
    return ~u64(0);
}

// count occurrences of every key of map `map`
__global__ void k_count_keys(const u64* sigs, u64 n, int h, int map, u64* tkeys, u32* tcount, u64 tmask,
                             int* overflow) {
    for (u64 r = blockIdx.x * u64(blockDim.x) + threadIdx.x; r < n; r += u64(gridDim.x) * blockDim.x) {
        const u64 key = sigs[r * h + map];
        const u64 slot = find_or_insert(tkeys, tmask, key);
        if (slot == ~u64(0)) {
            atomicExch(overflow, 1);
            continue;
        }
        atomicAdd(&tcount[slot], 1u);
    }
}

// assign value offsets to keys with 2..cap values; counters[0] = values,
// counters[1] = valid keys
__global__ void k_assign_offsets(const u64* tkeys, const u32* tcount, u64 tsize, int cap, u32* toff,
                                 unsigned long long* counters) {
    const int lane = lane_id();
    for (u64 base = blockIdx.x * u64(blockDim.x); base < tsize; base += u64(gridDim.x) * blockDim.x) {
        const u64 s = base + threadIdx.x;
        u32 cnt = 0;
        if (s < tsize && tkeys[s] != kEmptyKey) {
            const u32 c = tcount[s];
            if (c >= 2 && c <= u32(cap)) cnt = c;
        }
        // wave-aggregated offset allocation
        u32 incl = cnt;
        for (int o = 1; o < kWave; o <<= 1) {
            const u32 v = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += v;
        }
        const u32 waveTotal = __shfl(incl, kWave - 1, kWave);
        const unsigned long long valid = __ballot(cnt > 0);
        unsigned long long waveBase = 0;
        if (lane == 0 && waveTotal > 0) {
            waveBase = atomicAdd(&counters[0], (unsigned long long)waveTotal);
            atomicAdd(&counters[1], (unsigned long long)__popcll(valid));
        }
        waveBase = __shfl(waveBase, 0, kWave);
        if (s < tsize) toff[s] = cnt > 0 ? u32(waveBase + incl - cnt) : 0xFFFFFFFFu;
    }
}

// insert valid keys into the compact table
__global__ void k_insert_final(const u64* tkeys, const u32* tcount, const u32* toff, u64 tsize, u64* fkeys,
                               uint2* foffcnt, u64 fmask, int* overflow) {
    for (u64 s = blockIdx.x * u64(blockDim.x) + threadIdx.x; s < tsize; s += u64(gridDim.x) * blockDim.x) {
        if (toff[s] == 0xFFFFFFFFu) continue;
        const u64 key = tkeys[s];
        const u64 slot = find_or_insert(fkeys, fmask, key);
        if (slot == ~u64(0)) {
            atomicExch(overflow, 1);
            continue;
        }
        foffcnt[slot] = make_uint2(toff[s], tcount[s]);
    }
}

// scatter read ids into their key's value range
__global__ void k_scatter_values(const u64* sigs, u64 n, int h, int map, const u64* tkeys, const u32* toff, u64 tmask,
                                 u32* tfill, u32* values) {
    for (u64 r = blockIdx.x * u64(blockDim.x) + threadIdx.x; r < n; r += u64(gridDim.x) * blockDim.x) {
        const u64 key = sigs[r * h + map];
        const u64 slot = find_slot(tkeys, tmask, key);
        if (slot == ~u64(0)) continue;
        const u32 off = toff[slot];
        if (off == 0xFFFFFFFFu) continue;
        const u32 pos = atomicAdd(&tfill[slot], 1u);
        values[u64(off) + pos] = u32(r);
    }
}

// ascending read ids within every bucket (insertion sort, buckets <= cap)
__global__ void k_sort_buckets(const u32* toff, const u32* tcount, u64 tsize, u32* values) {
    for (u64 s = blockIdx.x * u64(blockDim.x) + threadIdx.x; s < tsize; s += u64(gridDim.x) * blockDim.x) {
        const u32 off = toff[s];
        if (off == 0xFFFFFFFFu) continue;
        u32* v = values + off;
        const int c = int(tcount[s]);
        for (int i = 1; i < c; i++) {
            const u32 x = v[i];
            int j = i - 1;
            while (j >= 0 && v[j] > x) {
                v[j + 1] = v[j];
                j--;
            }
            v[j + 1] = x;
        }
    }
}

static inline int grid_for(u64 n, int threads) {
    const u64 b = (n + threads - 1) / threads;
    return int(b < 65536 ? (b == 0 ? 1 : b) : 65536);
}

void launch_fill_u64(u64* p, u64 n, u64 v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_u32(u32* p, u64 n, u32 v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_count_keys(const u64* sigs, u64 n, int h, int map, u64* tkeys, u32* tcount, u64 tmask, int* overflow,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_count_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, sigs, n, h, map, tkeys, tcount, tmask,
                       overflow);
}
void launch_assign_offsets(const u64* tkeys, const u32* tcount, u64 tsize, int cap, u32* toff,
                           unsigned long long* counters, hipStream_t s) {
    hipLaunchKernelGGL(k_assign_offsets, dim3(grid_for(tsize, 256)), dim3(256), 0, s, tkeys, tcount, tsize, cap, toff,
                       counters);
}
void launch_insert_final(const u64* tkeys, const u32* tcount, const u32* toff, u64 tsize, u64* fkeys, uint2* foffcnt,
                         u64 fmask, int* overflow, hipStream_t s) {
    hipLaunchKernelGGL(k_insert_final, dim3(grid_for(tsize, 256)), dim3(256), 0, s, tkeys, tcount, toff, tsize, fkeys,
                       foffcnt, fmask, overflow);
}
void launch_scatter_values(const u64* sigs, u64 n, int h, int map, const u64* tkeys, const u32* toff, u64 tmask,
                           u32* tfill, u32* values, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_values, dim3(grid_for(n, 256)), dim3(256), 0, s, sigs, n, h, map, tkeys, toff, tmask,
                       tfill, values);
}
void launch_sort_buckets(const u32* toff, const u32* tcount, u64 tsize, u32* values, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_buckets, dim3(grid_for(tsize, 256)), dim3(256), 0, s, toff, tcount, tsize, values);
}

// ---------------------------------------------------------------------------
// Query: one block per anchor.  Signature, probes of all maps, gather of
// the hit lists into LDS, bitonic sort, unique, self/ambiguous removal.
// ---------------------------------------------------------------------------
// With candLimit > 0 only the candLimit candidates hit by the most maps are
// kept (ties broken by smaller read id), still in ascending id order.
__global__ void k_query(DevReadStore rs, DevTables t, u64 first, int n, int maxCand, int sortCap, int excludeAmbiguous,
                        int candLimit, u32* candIds, int* candCount, unsigned long long* rawHits) {
    HIP_DYNAMIC_SHARED(u64, lds64)
    u64* canon = lds64;                                  // maxLen
    u64* partial = canon + rs.maxLen;                    // blockDim
    u64* sig = partial + blockDim.x;                     // h
    int* counts = reinterpret_cast<int*>(sig + t.h);     // h + 1 (prefix)
    int* scanScratch = counts + t.h + 2;                 // 17
    u32* ids = reinterpret_cast<u32*>(scanScratch + 18); // sortCap
    u8* hits = reinterpret_cast<u8*>(ids + sortCap);      // maxCand (candLimit > 0 only)
    int* hist = counts;                                  // h + 1 bins, reused after the gather
    __shared__ int sTotal, sThresh, sNeed;

    for (int a = blockIdx.x; a < n; a += gridDim.x) {
        const u64 id = first + u64(a);
        const int len = rs.lens[id];
        const bool skip = (excludeAmbiguous && rs.ambig[id]) || len < t.k;
        if (skip) {
            if (threadIdx.x == 0) candCount[a] = 0;
            continue;  // uniform per block
        }
        if (t.sigs) {
            // signatures computed once for the table build
            for (int j = threadIdx.x; j < t.h; j += blockDim.x) sig[j] = t.sigs[id * u64(t.h) + j];
            __syncthreads();
        } else {
            block_signature(rs.seqs + id * rs.pitch, len, t.k, t.h, canon, partial, sig);
        }
        // probe every map
        for (int j = threadIdx.x; j < t.h; j += blockDim.x) {
            const DevMap m = t.maps[j];
            const u64 slot = find_slot(m.keys, m.mask, sig[j]);
            int c = 0;
            if (slot != ~u64(0)) {
                const uint2 oc = m.offcnt[slot];
                c = int(oc.y);
                partial[j] = (u64(oc.x) << 32) | u64(oc.y);  // reuse partial as (offset,count)
            } else {
                partial[j] = 0;
            }
            counts[j] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int j = 0; j < t.h; j++) {
                const int c = counts[j];
                counts[j] = run;
                run += c;
            }
            counts[t.h] = run;
            sTotal = run;
        }
        __syncthreads();
        const int total = sTotal;
        if (rawHits && threadIdx.x == 0) atomicAdd(rawHits, (unsigned long long)total);
        // gather (total <= h * cap <= sortCap by construction)
        const int n2 = int(next_pow2(unsigned(total > 0 ? total : 1)));
        for (int i = threadIdx.x; i < n2; i += blockDim.x) {
            u32 v = 0xFFFFFFFFu;
            if (i < total) {
                int lo = 0, hi = t.h - 1;  // last j with counts[j] <= i
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (counts[mid] <= i) lo = mid;
                    else hi = mid - 1;
                }
                const u64 oc = partial[lo];
                const u32 off = u32(oc >> 32);
                v = t.maps[lo].values[u64(off) + u64(i - counts[lo])];
            }
            ids[i] = v;
        }
        __syncthreads();
        // bitonic sort of n2 entries
        for (int size = 2; size <= n2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = threadIdx.x; i < n2 / 2; i += blockDim.x) {
                    const int lo = 2 * i - (i & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const u32 x = ids[lo], y = ids[hi];
                    if ((x > y) == up) {
                        ids[lo] = y;
                        ids[hi] = x;
                    }
                }
                __syncthreads();
            }
        }
        // unique, drop self and (optionally) ambiguous reads, compact in order
        u32* out = candIds + u64(a) * maxCand;
        int written = 0;
        for (int base = 0; base < total; base += blockDim.x) {
            const int i = base + threadIdx.x;
            bool keep = false;
            u32 v = 0;
            if (i < total) {
                v = ids[i];
                keep = (i == 0 || ids[i - 1] != v) && v != u32(id);
                if (keep && excludeAmbiguous) keep = rs.ambig[v] == 0;
            }
            int cnt;
            const int pos = block_exclusive_count(keep, scanScratch, &cnt);
            if (keep && written + pos < maxCand) {
                out[written + pos] = v;
                if (candLimit > 0) {
                    int run = 1;  // maps that hit v: length of its run in the sorted hits
                    while (i + run < total && ids[i + run] == v) run++;
                    hits[written + pos] = u8(run);
                }
            }
            written += cnt;
        }
        if (written > maxCand) written = maxCand;
        if (candLimit > 0 && written > candLimit) {
            // threshold hit count T: keep every candidate above T and the
            // first (in id order) candidates at T up to candLimit
            __syncthreads();
            for (int b = threadIdx.x; b <= t.h; b += blockDim.x) hist[b] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < written; i += blockDim.x) atomicAdd(&hist[hits[i]], 1);
            __syncthreads();
            if (threadIdx.x == 0) {
                int above = 0, v = t.h;
                while (v > 0 && above + hist[v] < candLimit) above += hist[v--];
                sThresh = v;
                sNeed = candLimit - above;
            }
            __syncthreads();
            const int T = sThresh, need = sNeed;
            int kept = 0, atT = 0;
            for (int base = 0; base < written; base += blockDim.x) {
                const int i = base + threadIdx.x;
                u32 v = 0;
                int hc = 0;
                if (i < written) {
                    v = out[i];
                    hc = hits[i];
                }
                int nT;
                const int rT = block_exclusive_count(i < written && hc == T, scanScratch, &nT);
                const bool keep = i < written && (hc > T || (hc == T && atT + rT < need));
                int cnt;
                const int pos = block_exclusive_count(keep, scanScratch, &cnt);
                if (keep) out[kept + pos] = v;
                kept += cnt;
                atT += nT;
            }
            written = kept;
        }
        if (threadIdx.x == 0) {
            candCount[a] = written;
            if (rawHits) atomicAdd(rawHits + 1, (unsigned long long)written);
        }
        __syncthreads();
    }
}

int query_lds_bytes(int h, int cap, int maxLen, int threads, int candLimit) {
    const int sortCap = int(1u << (32 - __builtin_clz(unsigned(h * cap > 1 ? h * cap - 1 : 1))));
    return int(sizeof(u64) * (size_t(maxLen) + threads + h) + sizeof(int) * (h + 2 + 18) + sizeof(u32) * sortCap +
               (candLimit > 0 ? size_t(h) * cap : 0));
}

void launch_query(const DevReadStore& rs, const DevTables& t, u64 first, int n, int maxCand, int excludeAmbiguous,
                  int candLimit, u32* candIds, int* candCount, unsigned long long* rawHits, hipStream_t stream) {
    if (n <= 0) return;
    const int threads = 256;
    const int lds = query_lds_bytes(t.h, maxCand / t.h, rs.maxLen, threads, candLimit);
    const int sortCap = int(1u << (32 - __builtin_clz(unsigned(maxCand > 1 ? maxCand - 1 : 1))));
    hipLaunchKernelGGL(k_query, dim3(n), dim3(threads), lds, stream, rs, t, first, n, maxCand, sortCap,
                       excludeAmbiguous, candLimit, candIds, candCount, rawHits);
}

}  // namespace care_engine
