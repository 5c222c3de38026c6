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
        // lengths beyond the LDS row (a caller's sequences) are cut to it
        block_signature(rs.seqs + id * rs.pitch, min(rs.lens[id], rs.maxLen), k, h, canon, partial, sig);
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

// Value offsets of the keys with 2..cap values, in slot order: every block
// sums a tile of kTile slots (values, keys), one block scans the tile sums,
// then every block assigns the offsets of its tile.  Deterministic, and no
// same-address atomics.  counters[0] = values, counters[1] = valid keys.
constexpr int kTile = 4096;

__device__ __forceinline__ u32 slot_values(const u64* tkeys, const u32* tcount, u64 s, u64 tsize, int cap) {
    if (s >= tsize || tkeys[s] == kEmptyKey) return 0;
    const u32 c = tcount[s];
    return (c >= 2 && c <= u32(cap)) ? c : 0;
}

// exclusive prefix sum of v over the block in thread order; *total gets the
// block's sum.  scratch: blockDim / 64 + 1 u64.  All threads must call.
__device__ __forceinline__ u64 block_exclusive_sum(u64 v, u64* scratch, u64* total) {
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int nwaves = (blockDim.x + kWave - 1) / kWave;
    u64 incl = v;
    for (int o = 1; o < kWave; o <<= 1) {
        const u64 t = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += t;
    }
    __syncthreads();
    if (lane == kWave - 1) scratch[wave] = incl;
    __syncthreads();
    u64 before = 0, all = 0;
    for (int w = 0; w < nwaves; w++) {
        const u64 c = scratch[w];
        if (w < wave) before += c;
        all += c;
    }
    *total = all;
    return before + incl - v;
}

__global__ void __launch_bounds__(256) k_tile_sums(const u64* tkeys, const u32* tcount, u64 tsize, int cap,
                                                   u64* tileSums) {
    __shared__ u64 scr[8];
    const u64 t0 = u64(blockIdx.x) * kTile;
    u64 vals = 0, keys = 0;
    for (int i = threadIdx.x; i < kTile; i += blockDim.x) {
        const u32 c = slot_values(tkeys, tcount, t0 + u64(i), tsize, cap);
        vals += c;
        keys += c > 0;
    }
    u64 tv, tk;
    block_exclusive_sum(vals, scr, &tv);
    block_exclusive_sum(keys, scr, &tk);
    if (threadIdx.x == 0) {
        tileSums[2 * u64(blockIdx.x)] = tv;
        tileSums[2 * u64(blockIdx.x) + 1] = tk;
    }
}

// one block: tileSums[2t] becomes the first value offset of tile t
__global__ void __launch_bounds__(1024) k_scan_tiles(u64* tileSums, u64 nt, unsigned long long* counters) {
    __shared__ u64 scr[17];
    u64 carryV = 0, carryK = 0;
    for (u64 b = 0; b < nt; b += blockDim.x) {
        const u64 t = b + threadIdx.x;
        const u64 v = t < nt ? tileSums[2 * t] : 0;
        const u64 k = t < nt ? tileSums[2 * t + 1] : 0;
        u64 sv, sk;
        const u64 ev = block_exclusive_sum(v, scr, &sv);
        block_exclusive_sum(k, scr, &sk);
        if (t < nt) tileSums[2 * t] = carryV + ev;
        carryV += sv;
        carryK += sk;
    }
    if (threadIdx.x == 0) {
        counters[0] = carryV;
        counters[1] = carryK;
    }
}

__global__ void __launch_bounds__(256) k_tile_offsets(const u64* tkeys, const u32* tcount, u64 tsize, int cap,
                                                      const u64* tileSums, u32* toff) {
    __shared__ u64 scr[8];
    const u64 t0 = u64(blockIdx.x) * kTile;
    u64 run = tileSums[2 * u64(blockIdx.x)];
    for (int r = 0; r < kTile; r += blockDim.x) {
        const u64 s = t0 + u64(r) + threadIdx.x;
        const u32 c = slot_values(tkeys, tcount, s, tsize, cap);
        u64 sum;
        const u64 ex = block_exclusive_sum(c, scr, &sum);
        if (s < tsize) toff[s] = c > 0 ? u32(run + ex) : 0xFFFFFFFFu;
        run += sum;
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

// (key, offset, count) triples of a prebuilt map into the compact table
__global__ void k_insert_triples(const u64* keys, const u32* offs, const u32* counts, u64 numKeys, u64* fkeys,
                                 uint2* foffcnt, u64 fmask, int* overflow) {
    for (u64 i = blockIdx.x * u64(blockDim.x) + threadIdx.x; i < numKeys; i += u64(gridDim.x) * blockDim.x) {
        const u64 slot = find_or_insert(fkeys, fmask, keys[i]);
        if (slot == ~u64(0)) {
            atomicExch(overflow, 1);
            continue;
        }
        foffcnt[slot] = make_uint2(offs[i], counts[i]);
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
u64 assign_offsets_scratch(u64 tsize) { return 2 * ((tsize + kTile - 1) / kTile); }

void launch_assign_offsets(const u64* tkeys, const u32* tcount, u64 tsize, int cap, u32* toff, u64* tileSums,
                           unsigned long long* counters, hipStream_t s) {
    const u64 nt = (tsize + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_tile_sums, dim3(unsigned(nt)), dim3(256), 0, s, tkeys, tcount, tsize, cap, tileSums);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tileSums, nt, counters);
    hipLaunchKernelGGL(k_tile_offsets, dim3(unsigned(nt)), dim3(256), 0, s, tkeys, tcount, tsize, cap, tileSums, toff);
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
void launch_insert_triples(const u64* keys, const u32* offs, const u32* counts, u64 numKeys, u64* fkeys,
                           uint2* foffcnt, u64 fmask, int* overflow, hipStream_t s) {
    hipLaunchKernelGGL(k_insert_triples, dim3(grid_for(numKeys, 256)), dim3(256), 0, s, keys, offs, counts, numKeys,
                       fkeys, foffcnt, fmask, overflow);
}

// ---------------------------------------------------------------------------
// Query: one wavefront per anchor.  Lane j probes map j with the anchor's
// stored signature value j, the h hit lists (each sorted ascending) are
// gathered and merged pairwise, then unique, self and ambiguous reads are
// removed in order.  Hit lists of up to `capL` entries are merged in the
// wave's LDS slice; longer ones (heavy repeats) in the anchor's rows of the
// global scratch `spill` (4 * maxCand words per anchor).
// ---------------------------------------------------------------------------
// With candLimit > 0 only the candLimit candidates hit by the most maps are
// kept (ties broken by smaller read id), still in ascending id order.
namespace {

constexpr int kQueryHeader = 800;  // per wave: counts[h + 1] and list pointers[h]

__device__ __forceinline__ void wave_sync_mem() {
    // orders the wave's LDS and global writes before other lanes' reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ u64 lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

constexpr int kMergeIlp = 4;  // entries per lane searched together (independent LDS chains)

// gathers the hit lists into a and merges them; returns the buffer (a or b)
// that holds the sorted entries.  Round r merges adjacent runs of 2^r lists
// (left run [l0, mid), right run [mid, r1)): an entry's merged position is
// its rank in its own run plus the entries of the other run before it
// (left: strictly smaller, right: smaller or equal), a stable merge.  Both
// searches (the entry's merged run, its rank in the other run) are
// fixed-step binary liftings over kMergeIlp entries at a time, so a lane
// has kMergeIlp independent LDS reads in flight instead of one chain.
template <class Buf>
__device__ __forceinline__ Buf gather_merge(Buf a, Buf b, const int* counts, const u32* const* lists, int h,
                                            int total, int lane) {
    for (int i = lane; i < total; i += kWave) {
        int lo = 0, hi = h - 1;  // last list j with counts[j] <= i
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (counts[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        a[i] = lists[lo][i - counts[lo]];
    }
    wave_sync_mem();
    Buf src = a, dst = b;
    for (int r = 0; (1 << r) < h; r++) {
        const int span = 2 << r;  // lists per merged run
        const int nseg = (h + span - 1) / span;
        // longest merged run of this round (bounds the rank search)
        int longest = lane < nseg ? counts[min(lane * span + span, h)] - counts[lane * span] : 0;
        for (int o = 32; o > 0; o >>= 1) longest = max(longest, __shfl_xor(longest, o, kWave));
        int segTop = 1, runTop = 1;
        while (2 * segTop <= nseg - 1) segTop *= 2;
        while (2 * runTop <= longest) runTop *= 2;
        for (int base = 0; base < total; base += kMergeIlp * kWave) {
            u32 x[kMergeIlp];
            int sg[kMergeIlp], rank[kMergeIlp], ob[kMergeIlp], olen[kMergeIlp], start[kMergeIlp], l0[kMergeIlp];
            bool left[kMergeIlp];
#pragma unroll
            for (int e = 0; e < kMergeIlp; e++) {
                const int i = base + e * kWave + lane;
                x[e] = i < total ? src[i] : 0u;
                sg[e] = 0;
            }
            // last merged run starting at or before i
            for (int step = segTop; step > 0; step >>= 1) {
#pragma unroll
                for (int e = 0; e < kMergeIlp; e++) {
                    const int i = base + e * kWave + lane;
                    const int t = sg[e] + step;
                    if (t < nseg && counts[t * span] <= i) sg[e] = t;
                }
            }
#pragma unroll
            for (int e = 0; e < kMergeIlp; e++) {
                const int i = base + e * kWave + lane;
                const int p0 = sg[e] * span;
                l0[e] = counts[p0];
                const int mid = counts[min(p0 + (span >> 1), h)];
                const int r1 = counts[min(p0 + span, h)];
                left[e] = i < mid;
                ob[e] = left[e] ? mid : l0[e];
                olen[e] = left[e] ? r1 - mid : mid - l0[e];
                start[e] = left[e] ? l0[e] : mid;
                rank[e] = 0;
            }
            // entries of the other run before x: a prefix of that run
            for (int step = runTop; step > 0; step >>= 1) {
#pragma unroll
                for (int e = 0; e < kMergeIlp; e++) {
                    const int t = rank[e] + step;
                    if (t <= olen[e]) {
                        const u32 v = src[ob[e] + t - 1];
                        if (left[e] ? v < x[e] : v <= x[e]) rank[e] = t;
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < kMergeIlp; e++) {
                const int i = base + e * kWave + lane;
                if (i < total) dst[l0[e] + (i - start[e]) + rank[e]] = x[e];
            }
        }
        wave_sync_mem();
        const Buf t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Bitonic sort of 64*K entries held in registers, entry e = lane*K + r in
// register r: strides below K stay inside a lane, larger ones exchange with
// lane ^ (stride / K).  Every loop bound is a constant, so the network is
// fully unrolled and v stays in registers.
template <int K>
__device__ __forceinline__ void sort_in_registers(u32 (&v)[K], int lane) {
#pragma unroll
    for (int size = 2; size <= kWave * K; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= K) {
                const int lx = stride / K;
#pragma unroll
                for (int r = 0; r < K; r++) {
                    const u32 o = __shfl_xor(v[r], lx, kWave);
                    const int e = lane * K + r;
                    const bool up = (e & size) == 0;
                    const bool lower = (e & stride) == 0;
                    v[r] = (lower == up) ? min(v[r], o) : max(v[r], o);
                }
            } else {
#pragma unroll
                for (int r = 0; r < K; r++) {
                    if (r & stride) continue;
                    const int e = lane * K + r;
                    const bool up = (e & size) == 0;
                    const u32 a = v[r], b = v[r + stride];
                    v[r] = up ? min(a, b) : max(a, b);
                    v[r + stride] = up ? max(a, b) : min(a, b);
                }
            }
        }
    }
}

// gathers the hit lists (total <= 64*K) straight into registers, padded
// with ~0u (sorts last), sorts them and stores the sorted entries to out
template <int K, class Buf>
__device__ __forceinline__ void gather_sort(Buf out, const int* counts, const u32* const* lists, int h, int total,
                                            int lane) {
    // a lane's K entries are consecutive: one binary search finds the list
    // of its first entry, the others step through the following lists
    // (counts[h] = total ends the walk)
    const int e0 = lane * K;
    int j = 0;
    if (e0 < total) {
        int hi = h - 1;  // last list j with counts[j] <= e0
        while (j < hi) {
            const int mid = (j + hi + 1) >> 1;
            if (counts[mid] <= e0) j = mid;
            else hi = mid - 1;
        }
    }
    int start = counts[j], next = counts[j + 1];
    const u32* list = lists[j];
    u32 v[K];
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int e = e0 + r;
        u32 x = 0xFFFFFFFFu;
        if (e < total) {
            while (e >= next) {
                j++;
                start = next;
                next = counts[j + 1];
                list = lists[j];
            }
            x = list[e - start];
        }
        v[r] = x;
    }
    sort_in_registers<K>(v, lane);
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int e = lane * K + r;
        if (e < total) out[e] = v[r];
    }
    wave_sync_mem();
}

// longest hit list sorted in registers: 16 per lane in the kernel for tables
// whose lists stay short (h * cap <= 1024), 32 per lane otherwise (more
// registers, fewer lists left to the merge); longer lists are merged
constexpr int kRegisterSortMax = 32 * kWave;

template <int kMaxK, class Buf>
__device__ __forceinline__ void gather_sort_any(Buf out, const int* counts, const u32* const* lists, int h, int total,
                                                int lane) {
    if (total <= 2 * kWave) gather_sort<2>(out, counts, lists, h, total, lane);
    else if (total <= 4 * kWave) gather_sort<4>(out, counts, lists, h, total, lane);
    else if (total <= 8 * kWave) gather_sort<8>(out, counts, lists, h, total, lane);
    else if (kMaxK == 16 || total <= 16 * kWave) gather_sort<16>(out, counts, lists, h, total, lane);
    else gather_sort<(kMaxK > 16 ? kMaxK : 16)>(out, counts, lists, h, total, lane);
}

// unique, self/ambiguous removal and compaction into out; hit counts
// (lengths of the runs of equal ids) into hits when candLimit > 0
template <class Buf, class HitBuf>
__device__ __forceinline__ int compact_unique(Buf sorted, int total, u32 self, const DevReadStore& rs,
                                              int excludeAmbiguous, int candLimit, u32* out, HitBuf hits, int lane) {
    int written = 0;
    const u64 below = lanes_below(lane);
    // hit counts (candLimit > 0): the maps that hit v are the length of its
    // run of equal entries (at most h <= 64), which ends where the next run
    // starts; the last kept run of a chunk waits in (pendPos, pendStart)
    // until a later chunk's first run start, or the end of the list
    int pendPos = -1, pendStart = 0;
    for (int base = 0; base < total; base += kWave) {
        const int i = base + lane;
        bool start = false, keep = false;
        u32 v = 0;
        if (i < total) {
            v = sorted[i];
            start = i == 0 || sorted[i - 1] != v;
            keep = start && v != self;
            if (keep && excludeAmbiguous) keep = rs.ambig[v] == 0;
        }
        const u64 bal = __ballot(keep);
        const int pos = written + __popcll(bal & below);
        if (keep) {
            CARE_CHECK(v < rs.n && pos < total);
            out[pos] = v;
        }
        if (candLimit > 0) {
            const u64 starts = __ballot(start);  // the same in every lane
            if (starts) {
                if (pendPos >= 0) {
                    if (lane == 0) hits[pendPos] = u8(base + __ffsll((long long)starts) - 1 - pendStart);
                    pendPos = -1;
                }
                const u64 after = starts & ~(below | (1ull << lane));  // run starts after this lane's
                if (keep && after) hits[pos] = u8(base + __ffsll((long long)after) - 1 - i);
                const int last = 63 - __clzll((long long)starts);
                if ((bal >> last) & 1ull) {
                    pendPos = written + __popcll(bal & ((1ull << last) - 1ull));
                    pendStart = base + last;
                }
            }
        }
        written += __popcll(bal);
    }
    if (pendPos >= 0 && lane == 0) hits[pendPos] = u8(total - pendStart);
    return written;
}

// keeps the candLimit candidates with the most hits (ties: smaller id, i.e.
// earlier in out), in place and in order; hist holds h + 1 ints
template <class HitBuf>
__device__ __forceinline__ int keep_top(u32* out, HitBuf hits, int written, int candLimit, int h, int* hist,
                                        int lane) {
    for (int b = lane; b <= h; b += kWave) hist[b] = 0;
    wave_sync_mem();
    for (int i = lane; i < written; i += kWave) atomicAdd(&hist[hits[i]], 1);
    wave_sync_mem();
    // threshold T: the largest hit count at which the candidates with at
    // least T hits reach candLimit (0 if they never do); lane l holds the
    // count of hit number h - l, prefix sums run from h downwards
    const int cnt = lane < h ? hist[h - lane] : 0;
    int incl = cnt;
    for (int o = 1; o < kWave; o <<= 1) {
        const int v = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += v;
    }
    const u64 reach = __ballot(lane < h && incl >= candLimit);
    int T, need;
    if (reach) {
        const int l = __ffsll((long long)reach) - 1;
        T = h - l;
        need = candLimit - (__shfl(incl, l, kWave) - __shfl(cnt, l, kWave));
    } else {
        T = 0;
        need = candLimit - __shfl(incl, h - 1, kWave);
    }
    const u64 below = lanes_below(lane);
    int kept = 0, atT = 0;
    for (int base = 0; base < written; base += kWave) {
        const int i = base + lane;
        u32 v = 0;
        int hc = 0;
        if (i < written) {
            v = out[i];
            hc = hits[i];
        }
        const u64 bT = __ballot(i < written && hc == T);
        const int rT = __popcll(bT & below);
        const bool keep = i < written && (hc > T || (hc == T && atT + rT < need));
        const u64 bk = __ballot(keep);
        if (keep) out[kept + __popcll(bk & below)] = v;  // kept + rank <= i: reads of later entries unaffected
        kept += __popcll(bk);
        atT += __popcll(bT);
    }
    return kept;
}

}  // namespace

template <int kMaxK>
__global__ void __launch_bounds__(256) k_query(DevReadStore rs, DevTables t, u64 first, int n, int maxCand, int capL,
                                               int regMax, int excludeAmbiguous, int candLimit, u32* candIds,
                                               int* candCount, u32* spill, int* rawCount, QueryAnchors qa) {
    HIP_DYNAMIC_SHARED(uint4, ldsRaw)
    const int wave = threadIdx.x / kWave;
    const int wpb = blockDim.x / kWave;
    const int h = t.h;
    char* wl = reinterpret_cast<char*>(ldsRaw) + size_t(wave) * query_wave_lds_bytes(capL, candLimit);
    int* counts = reinterpret_cast<int*>(wl);                       // h + 1, reused as hit histogram
    const u32** lists = reinterpret_cast<const u32**>(wl + 288);     // h
    u32* la = reinterpret_cast<u32*>(wl + kQueryHeader);            // capL
    u32* lb = la + capL;                                             // capL
    u8* lh = reinterpret_cast<u8*>(lb + capL);                       // capL (candLimit > 0)
#ifdef CARE_EMU_CHECKS
    CARE_CHECK(size_t(wpb) * query_wave_lds_bytes(capL, candLimit) <= emu::dyn_lds_bytes() && h + 1 <= 72);
#endif

    for (int a = blockIdx.x * wpb + wave; a < n; a += gridDim.x * wpb) {
        // anchor read id (~0 for a caller's sequence without one), length
        // and signature row
        const u64 id = qa.ids ? u64(qa.ids[a]) : qa.sigs ? ~u64(0) : first + u64(a);
        const bool stored = id < rs.n;
        // the lane index is made opaque in every iteration: otherwise the
        // compiler hoists the lane-dependent addresses of the sort and merge
        // networks out of this loop and keeps them all in registers (about
        // 190 instead of 104 VGPRs for the 32-per-lane sort)
        int lane = lane_id();
        asm volatile("" : "+v"(lane));
        const int len = qa.lens ? qa.lens[a] : rs.lens[id];
        const u64* sig = qa.sigs ? qa.sigs + u64(a) * u64(h) : t.sigs + id * u64(h);
        const bool skip = (excludeAmbiguous && stored && rs.ambig[id]) || len < t.k;
        if (skip) {
            if (lane == 0) {
                candCount[a] = 0;
                if (rawCount) rawCount[a] = 0;
            }
            continue;  // uniform per wave
        }
        // lane j probes map j with the stored signature value j
        int c = 0;
        const u32* list = nullptr;
        if (lane < h) {
            const DevMap m = t.maps[lane];
            const u64 slot = find_slot(m.keys, m.mask, sig[lane]);
            if (slot != ~u64(0)) {
                const uint2 oc = m.offcnt[slot];
                c = int(oc.y);
                list = m.values + oc.x;
            }
        }
        int incl = c;
        for (int o = 1; o < kWave; o <<= 1) {
            const int v = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += v;
        }
        const int total = __shfl(incl, kWave - 1, kWave);
        CARE_CHECK(total <= maxCand && (stored || qa.sigs));
        if (lane < h) {
            counts[lane] = incl - c;
            lists[lane] = list;
        }
        if (lane == 0) counts[h] = total;
        wave_sync_mem();

        u32* out = candIds + u64(a) * maxCand;
        u32* gs = spill + u64(a) * 4 * maxCand;  // [maxCand] + [maxCand] + hit bytes
        u8* gh = reinterpret_cast<u8*>(gs + 2 * maxCand);
        int written;
        if (total <= capL) {
            const u32* sorted = la;
            if (total <= regMax) gather_sort_any<kMaxK>(la, counts, lists, h, total, lane);
            else sorted = gather_merge(la, lb, counts, lists, h, total, lane);
            written = compact_unique(sorted, total, u32(id), rs, excludeAmbiguous, candLimit, out, lh, lane);
        } else {
            const u32* sorted = gs;
            if (total <= regMax) gather_sort_any<kMaxK>(gs, counts, lists, h, total, lane);
            else sorted = gather_merge(gs, gs + maxCand, counts, lists, h, total, lane);
            written = compact_unique(sorted, total, u32(id), rs, excludeAmbiguous, candLimit, out, gh, lane);
        }
        wave_sync_mem();
        if (candLimit > 0 && written > candLimit) {
            written = total <= capL ? keep_top(out, lh, written, candLimit, h, counts, lane)
                                    : keep_top(out, gh, written, candLimit, h, counts, lane);
        }
        if (lane == 0) {
            candCount[a] = written;
            if (rawCount) rawCount[a] = total;
        }
        wave_sync_mem();  // counts and buffers are rewritten by the next anchor
    }
}

// totals[0] += sum of raw hit counts, totals[1] += sum of candidate counts
// (one workgroup, so no atomics)
__global__ void __launch_bounds__(1024) k_sum_counts(const int* rawCount, const int* candCount, int n,
                                                     unsigned long long* totals) {
    __shared__ unsigned long long part[2][16];
    unsigned long long r = 0, c = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        r += u64(rawCount[i]);
        c += u64(candCount[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        r += __shfl_xor(r, o, kWave);
        c += __shfl_xor(c, o, kWave);
    }
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        part[0][w] = r;
        part[1][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < int(blockDim.x / kWave); k++) {
            r += part[0][k];
            c += part[1][k];
        }
        totals[0] += r;
        totals[1] += c;
    }
}

int query_lds_bytes(int h, int cap, int maxLen, int threads, int candLimit, int capL) {
    (void)h;
    (void)cap;
    (void)maxLen;
    return int(size_t(threads / kWave) * query_wave_lds_bytes(capL, candLimit));
}

void launch_query(const DevReadStore& rs, const DevTables& t, u64 first, int n, int maxCand, int capL, int regMax,
                  int excludeAmbiguous, int candLimit, u32* candIds, int* candCount, u32* spill, int* rawCount,
                  hipStream_t stream, const QueryAnchors& qa) {
    if (n <= 0) return;
    const int threads = 256;
    const int wpb = threads / kWave;
    const int lds = query_lds_bytes(t.h, maxCand / t.h, rs.maxLen, threads, candLimit, capL);
    const int blocks = std::min((n + wpb - 1) / wpb, 1 << 16);
    const bool wide = maxCand > 16 * kWave;  // lists may exceed the 16-per-lane network
    const int kernelMax = wide ? 32 * kWave : 16 * kWave;
    regMax = regMax > 0 ? std::min(regMax, kernelMax) : kernelMax;
    // the 32-per-lane instantiation holds about 190 VGPRs (two waves per
    // SIMD), the 16-per-lane one 128 (four): the wider register sort is
    // used only when asked for lists of more than 1024 hits
    if (regMax > 16 * kWave)
        hipLaunchKernelGGL(k_query<32>, dim3(blocks), dim3(threads), lds, stream, rs, t, first, n, maxCand, capL, regMax,
                           excludeAmbiguous, candLimit, candIds, candCount, spill, rawCount, qa);
    else
        hipLaunchKernelGGL(k_query<16>, dim3(blocks), dim3(threads), lds, stream, rs, t, first, n, maxCand, capL, regMax,
                           excludeAmbiguous, candLimit, candIds, candCount, spill, rawCount, qa);
}

void launch_sum_counts(const int* rawCount, const int* candCount, int n, unsigned long long* totals, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(1024), 0, s, rawCount, candCount, n, totals);
}

}  // namespace care_engine
