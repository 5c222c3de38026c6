// Fused per-anchor correction kernel: shifted-Hamming alignment, orientation
// choice, mismatch-ratio / pair filters, MSA build, different-region
// refinement, MSA properties, anchor correction and candidate correction.
//
// One workgroup corrects one anchor at a time and pulls anchors from a
// device-side work queue (persistent kernel), so anchors with very different
// candidate counts balance across the chip.  Semantics follow the reference
// CPU corrector exactly (include/corrector.hpp:145-536, src/msa.cpp,
// src/cpu_alignment.cpp, include/alignmentorientation.hpp), which the GPU
// reference does not (it subtracts removed candidates during refinement,
// include/gpu/gpumsa.cuh:743-812).  Floating-point sums run in the CPU's
// order: per column, anchor first, then candidates in ascending index.

#include <algorithm>

#include "device_common.h"
#include "mlp_device.h"

namespace care_engine {

namespace {

__device__ __forceinline__ bool feq(float l, float r) {
    const float d = l - r < 0 ? -(l - r) : l - r;
    return d < 1e-5f;
}
__device__ __forceinline__ bool fleq(float l, float r) { return (l < r) || feq(l, r); }
__device__ __forceinline__ bool fgeq(float l, float r) { return (l > r) || feq(l, r); }

// 16-byte candidate records moved as one vector load / store
union MetaBits {
    CandMeta m;
    uint4 v;
};
__device__ __forceinline__ CandMeta load_meta(const CandMeta* p) {
    MetaBits b;
    b.v = *reinterpret_cast<const uint4*>(p);
    return b.m;
}
__device__ __forceinline__ void store_meta(CandMeta* p, const CandMeta& m) {
    MetaBits b;
    b.m = m;
    *reinterpret_cast<uint4*>(p) = b.v;
}

// pack the even bits (0,2,..,30) of x into the low 16 bits
__device__ __forceinline__ u32 even_bits(u32 x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

// Shifts still under their mismatch limit after the first 64 bases, queued
// per wavefront: (candidate << 11 | order index, mismatches forward | reverse
// << 16).  A slice holds one wavefront's worth.
constexpr int kSurvivors = 64;

// LDS of one workgroup.  A sequence is held as interleaved bit planes: word
// k is (hi bits, lo bits) of bases 32k..32k+31, MSB first.  Planes carry one
// leading and two trailing zero words: plane word i lives at index i + 1, so
// an unaligned 32-bit window starting at base d is always
// alignbit(p[e], p[e + 1]) with e = d / 32 + (d % 32 != 0).
struct Lds {
    CandMeta* mst;  // [chunk] candidate records of the current MSA chunk
    u32* rows;      // [chunk][pitch] their 2-bit rows
    u32* qrows;     // [chunk][qpitch] their binned quality rows
    float* qw;      // [256] quality weight per quality character
    u32* aw;        // anchor 2-bit words [pitch]
    uint2* ap;      // anchor planes [hw + 3]
    char* anchorQ;  // [maxLen]
    u32* keys;      // [chunk][2] alignment keys (forward, reverse complement) of the staged candidates,
                    // followed by the survivor slices of the waves [waves][2][kSurvivors]
    uint2* cpl;     // [chunk] candidate planes: forward, then reverse complement, 2 x [hw + 3]
    u8* cons;       // [cols]
    float* support;
    int* coverage;
    int* origCov;
    u16* cnt;       // [4][cols]
    float* wgt;     // [4][cols]
    int* scratch;   // [32]
    u16* dec;       // gated mode: decision columns (anchor positions) [maxLen]
    u8* gacc;       // gated mode: accepted flag per anchor position [maxLen]
    char* gscr;     // gated mode: feature tile and MLP scratch, overlays mst/rows/qrows
};

struct LdsDims {
    int pitch, qpitch, hw, maxLen, cols, waves, chunk, gated;
    __host__ __device__ int pw() const { return hw + 3; }
    __host__ __device__ size_t bytes() const {
        size_t b = 0;
        b += size_t(chunk) * (sizeof(CandMeta) + sizeof(u32) * (pitch + qpitch));
        b += sizeof(float) * 256;
        b += sizeof(u32) * pitch;
        b += sizeof(uint2) * pw();
        b += (size_t(maxLen) + 3) / 4 * 4;
        b += size_t(chunk) * 2 * sizeof(uint2) * pw();
        b += (size_t(cols) + 3) / 4 * 4;
        b += sizeof(float) * cols + 2 * sizeof(int) * cols;
        b += 4 * sizeof(u16) * cols + 4 * sizeof(float) * cols;
        b += 32 * sizeof(int);
        b += 2 * sizeof(u32) * size_t(chunk);
        b += 2 * sizeof(u32) * size_t(kSurvivors) * size_t(waves);  // survivor slices after the keys
        if (gated & 1) b += (sizeof(u16) + 1) * ((size_t(maxLen) + 15) / 16 * 16);
        return b;
    }
};

__device__ Lds carve(char* base, const LdsDims& d) {
    Lds l;
    char* p = base;
    // 16-byte aligned base: 16- and 8-byte records first
    l.mst = reinterpret_cast<CandMeta*>(p); p += sizeof(CandMeta) * d.chunk;
    l.ap = reinterpret_cast<uint2*>(p); p += sizeof(uint2) * d.pw();
    l.cpl = reinterpret_cast<uint2*>(p); p += size_t(d.chunk) * 2 * sizeof(uint2) * d.pw();
    l.rows = reinterpret_cast<u32*>(p); p += sizeof(u32) * size_t(d.chunk) * d.pitch;
    l.qrows = reinterpret_cast<u32*>(p); p += sizeof(u32) * size_t(d.chunk) * d.qpitch;
    l.qw = reinterpret_cast<float*>(p); p += sizeof(float) * 256;
    l.aw = reinterpret_cast<u32*>(p); p += sizeof(u32) * d.pitch;
    l.support = reinterpret_cast<float*>(p); p += sizeof(float) * d.cols;
    l.coverage = reinterpret_cast<int*>(p); p += sizeof(int) * d.cols;
    l.origCov = reinterpret_cast<int*>(p); p += sizeof(int) * d.cols;
    l.wgt = reinterpret_cast<float*>(p); p += 4 * sizeof(float) * d.cols;
    l.scratch = reinterpret_cast<int*>(p); p += 32 * sizeof(int);
    l.cnt = reinterpret_cast<u16*>(p); p += 4 * sizeof(u16) * d.cols;
    l.cons = reinterpret_cast<u8*>(p); p += (size_t(d.cols) + 3) / 4 * 4;
    l.anchorQ = reinterpret_cast<char*>(p); p += (size_t(d.maxLen) + 3) / 4 * 4;
    l.keys = reinterpret_cast<u32*>(p); p += 2 * sizeof(u32) * size_t(d.chunk);
    p += 2 * sizeof(u32) * size_t(kSurvivors) * size_t(d.waves);  // survivor slices
    l.gscr = reinterpret_cast<char*>(l.mst);
    if (d.gated & 1) {
        const size_t ml = (size_t(d.maxLen) + 15) / 16 * 16;
        l.dec = reinterpret_cast<u16*>(p); p += sizeof(u16) * ml;
        l.gacc = reinterpret_cast<u8*>(p);
    } else {
        l.dec = nullptr;
        l.gacc = nullptr;
    }
    CARE_CHECK(size_t(p - base) + ((d.gated & 1) ? (size_t(d.maxLen) + 15) / 16 * 16 : 0) <= d.bytes());
#ifdef CARE_EMU_CHECKS
    CARE_CHECK(d.bytes() <= emu::dyn_lds_bytes());
#endif
    return l;
}

struct Msa {
    int cols;
    int ab;  // first anchor column
};

// Builds the MSA of the anchor and candidates meta[0..m) into LDS
// (src/msa.cpp:9-173).  One thread per column; per column the weights are
// summed anchor first, then candidates in index order, as on the CPU.
// Candidate records, 2-bit rows and quality rows are staged through LDS in
// chunks of dims.chunk candidates.  Only the columns within `margin` of the
// anchor are computed: every reader of the MSA (refinement, properties,
// anchor correction and its features, candidate correction within
// new_columns of the anchor and the gate features two columns around a
// corrected read) stays inside them, and columns are independent of each
// other.  All threads must call.
// per-column sums of the MSA: base counts and weight sums, the weights
// added in candidate order as the CPU adds them
struct MsaSums {
    int cA, cC, cG, cT;
    float wA, wC, wG, wT;
};

// conditional adds of +0.0f leave the other sums bit-identical (no sum is
// ever -0.0f); base 4 adds nothing
__device__ __forceinline__ void msa_add(MsaSums& s, u32 b, float w) {
    s.cA += b == 0u;
    s.cC += b == 1u;
    s.cG += b == 2u;
    s.cT += b == 3u;
    s.wA += b == 0u ? w : 0.0f;
    s.wC += b == 1u ? w : 0.0f;
    s.wG += b == 2u ? w : 0.0f;
    s.wT += b == 3u ? w : 0.0f;
}

// adds the staged candidates 0..nc-1 to the sums of MSA column colRel
// (relative to the anchor's first base).  kQ is the quality width (0: no
// qualities, else 1, 2 or 8 bits per base); q4 holds the weights of the
// 1- and 2-bit bins.  The body has no branch: the orientation selects
// indices, not code paths.  kSkipMiss: a wavefront skips the candidates
// that miss all of its columns (measured faster in the plain kernel and
// slower in the gated one).  Left rolled: unrolled copies hold more loads in
// flight and raise the registers of the calling kernels.
template <int kQ, bool kSkipMiss>
__device__ __forceinline__ void msa_add_staged(const Lds& L, int nc, int colRel, int pitch, int qpitch, int maxLen,
                                               const float (&q4)[4], MsaSums& s) {
#pragma unroll 1
    for (int c = 0; c < nc; c++) {
        const CandMeta cm = L.mst[c];
        const int len = cm.len;
        const int p0 = colRel - int(cm.shift);
        const bool in = u32(p0) < u32(len);
        if (kSkipMiss && !__any(in)) continue;  // the candidate misses every column of this wavefront
        const int pos = in ? p0 : 0;
        const bool rc = cm.flag == 2;
        // reverse complement: base len-1-pos of the stored read, complemented
        const int i = rc ? len - 1 - pos : pos;
        const u32 word = L.rows[c * pitch + (i >> 4)];
        const u32 b = ((word >> (30 - 2 * (i & 15))) & 3u) ^ (rc ? 3u : 0u);
        float w = cm.u.weight;
        if (kQ != 0) {
            const int qi = rc ? maxLen - 1 - pos : pos;
            const u32* qr = L.qrows + c * qpitch;
            float qw;
            if (kQ == 8) {
                qw = L.qw[(qr[qi >> 2] >> (8 * (qi & 3))) & 0xFFu];
            } else if (kQ == 2) {
                const u32 v = (qr[qi >> 4] >> (30 - 2 * (qi & 15))) & 3u;
                qw = v == 0u ? q4[0] : v == 1u ? q4[1] : v == 2u ? q4[2] : q4[3];
            } else {
                qw = ((qr[qi >> 5] >> (31 - (qi & 31))) & 1u) ? q4[1] : q4[0];
            }
            w = cm.u.weight * qw;
        }
        msa_add(s, in ? b : 4u, w);
    }
}

template <bool kSkipMiss>
__device__ Msa build_msa(const CorrectParams& P, const Lds& L, const LdsDims& dims, const CandMeta* meta, int m,
                         int la, int margin, int* sMin, int* sMax) {
    if (threadIdx.x == 0) {
        *sMin = 0;
        *sMax = la;
    }
    __syncthreads();
    int lmin = 0, lmax = la;
    for (int c = threadIdx.x; c < m; c += blockDim.x) {
        const int s = meta[c].shift;
        lmin = min(lmin, s);
        lmax = max(lmax, s + int(meta[c].len));
    }
    // wave reductions first: one LDS atomic per wave instead of per thread
    lmin = wave_min_int(lmin);
    lmax = wave_max_int(lmax);
    if (lane_id() == 0) {
        atomicMin(sMin, lmin);
        atomicMax(sMax, lmax);
    }
    __syncthreads();
    Msa msa;
    msa.cols = *sMax - *sMin;
    msa.ab = -*sMin;
    CARE_CHECK(msa.cols <= dims.cols && msa.ab >= 0);
    const int ae = msa.ab + la;
    const DevReadStore& rs = P.rs;
    const bool useQ = P.useQuality != 0;
    const int cols = msa.cols;
    const int pitch = rs.pitch, qpitch = useQ ? rs.qpitch : 0;
    CARE_CHECK(pitch == dims.pitch && qpitch == dims.qpitch);
    // staging assignment: thread -> (row r0 + k * rowsPerStep, word w)
    // thread index opaque: the staging assignment is computed per build, not
    // hoisted out of the anchor loop into registers held by every phase
    int tx = int(threadIdx.x);
    asm volatile("" : "+v"(tx));
    const int rowsS = int(blockDim.x) / pitch;
    const int r0S = tx / pitch, wS = tx - r0S * pitch;
    const int rowsQ = qpitch ? int(blockDim.x) / qpitch : 0;
    const int r0Q = qpitch ? tx / qpitch : 0, wQ = tx - r0Q * qpitch;
    const int wb = max(0, msa.ab - margin), we = min(cols, ae + margin);
    // quality width of the candidate rows, and the weights of the bin
    // characters of 2-bit ('$', '(', '1', '9') and 1-bit ('(', '5') rows
    const int qmode = useQ ? rs.qbits : 0;
    float q4[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (qmode == 2) {
        q4[0] = L.qw['$'];
        q4[1] = L.qw['('];
        q4[2] = L.qw['1'];
        q4[3] = L.qw['9'];
    } else if (qmode == 1) {
        q4[0] = L.qw['('];
        q4[1] = L.qw['5'];
    }
    for (int cb = wb; cb < we; cb += blockDim.x) {
        const int col = cb + int(threadIdx.x);
        const bool active = col < we;
        MsaSums sums = {0, 0, 0, 0, 0.f, 0.f, 0.f, 0.f};
        u32 anchorBase = 0;
        const bool inAnchor = active && col >= msa.ab && col < ae;
        if (inAnchor) {
            const int i = col - msa.ab;
            anchorBase = base_at(L.aw, i);
            const float w = useQ ? L.qw[(unsigned char)L.anchorQ[i]] : 1.0f;
            msa_add(sums, anchorBase, 1.0f * w);
        }
        for (int c0 = 0; c0 < m; c0 += dims.chunk) {
            const int nc = min(dims.chunk, m - c0);
            __syncthreads();
            for (int i = threadIdx.x; i < nc; i += blockDim.x) {
                CARE_CHECK(c0 + i < P.maxCand && meta[c0 + i].id < rs.n);
                store_meta(L.mst + i, load_meta(meta + c0 + i));
            }
            if (r0S < rowsS)
                for (int c = r0S; c < nc; c += rowsS)
                    L.rows[c * pitch + wS] = rs.seqs[u64(meta[c0 + c].id) * pitch + wS];
            if (r0Q < rowsQ)
                for (int c = r0Q; c < nc; c += rowsQ)
                    L.qrows[c * qpitch + wQ] = rs.quals[u64(meta[c0 + c].id) * qpitch + wQ];
            __syncthreads();
            // a wave-uniform test: lanes past the window run the loop with
            // the others (harmlessly, their sums are dropped), which keeps
            // its counter scalar
            if (!__any(active)) continue;
            const int colRel = col - msa.ab;
            const int ncU = __builtin_amdgcn_readfirstlane(nc);  // uniform: a scalar loop counter
            if (qmode == 0) msa_add_staged<0, kSkipMiss>(L, ncU, colRel, pitch, qpitch, rs.maxLen, q4, sums);
            else if (qmode == 2) msa_add_staged<2, kSkipMiss>(L, ncU, colRel, pitch, qpitch, rs.maxLen, q4, sums);
            else if (qmode == 1) msa_add_staged<1, kSkipMiss>(L, ncU, colRel, pitch, qpitch, rs.maxLen, q4, sums);
            else msa_add_staged<8, kSkipMiss>(L, ncU, colRel, pitch, qpitch, rs.maxLen, q4, sums);
        }
        if (!active) continue;
        const int cA = sums.cA, cC = sums.cC, cG = sums.cG, cT = sums.cT;
        const float wA = sums.wA, wC = sums.wC, wG = sums.wG, wT = sums.wT;
        u8 cons = 0;
        float cw = wA;
        if (wC > cw) { cons = 1; cw = wC; }
        if (wG > cw) { cons = 2; cw = wG; }
        if (wT > cw) { cons = 3; cw = wT; }
        const float colw = wA + wC + wG + wT;
        L.cons[col] = cons;
        L.support[col] = cw / colw;
        L.coverage[col] = cA + cC + cG + cT;
        L.cnt[0 * cols + col] = u16(cA);
        L.cnt[1 * cols + col] = u16(cC);
        L.cnt[2 * cols + col] = u16(cG);
        L.cnt[3 * cols + col] = u16(cT);
        L.wgt[0 * cols + col] = wA;
        L.wgt[1 * cols + col] = wC;
        L.wgt[2 * cols + col] = wG;
        L.wgt[3 * cols + col] = wT;
        if (inAnchor)
            L.origCov[col] = anchorBase == 0u ? cA : anchorBase == 1u ? cC : anchorBase == 2u ? cG : cT;
    }
    __syncthreads();
    return msa;
}

// masks of the two words j, j + 1 of a lane with nw words whose last word
// keeps the bits of lastMask
__device__ __forceinline__ void word_masks(int j, int nw, u32 lastMask, u32* m1, u32* m2) {
    const u32 a = j + 1 < nw ? ~0u : lastMask;
    const u32 b = j + 2 < nw ? ~0u : lastMask;
    *m1 = j < nw ? a : 0u;
    *m2 = j + 1 < nw ? b : 0u;
}

struct AlignCtx {
    const uint2* ap;   // anchor planes
    const uint2* cpl;  // candidate planes of the chunk, 2 x pw per candidate
    const CandMeta* mst;
    u32* keys;         // [chunk][2] keys, combined by atomic minimum
    u32* survA;        // this wave's survivor slice
    u32* survB;
    int la, pw, minov;
    float maxErrorRate;
};

// Finishes the queued survivors of a wavefront (one per lane): the words
// from 64 on, then the keys into the candidates' minima.  All lanes call.
__device__ __forceinline__ void finish_survivors(const AlignCtx& A, int count) {
    const int lane = lane_id();
    const bool valid = lane < count;
    u32 ea = 0, eb = 0;
    if (valid) {
        ea = A.survA[lane];
        eb = A.survB[lane];
    }
    const int c = int(ea >> 11), u = int(ea & 2047u);
    int mf = int(eb & 0xFFFFu), mr = int(eb >> 16);
    const int lc = valid ? int(A.mst[c].len) : 0;
    const int la = A.la;
    const int nPos = max(0, la - A.minov + 1);
    const bool pos = u < nPos;
    const int s = pos ? u : -(u - nPos + 1);
    const int ov = valid ? (pos ? min(la - s, lc) : min(la, lc + s)) : 0;
    const int lim = int(float(ov) * A.maxErrorRate);
    const int nw = valid ? (ov + 31) >> 5 : 0;
    const int d = pos ? s : -s;
    const u32 sh = u32(32 - (d & 31)) & 31u;
    const int off = (d >> 5) + ((d & 31) != 0);
    const uint2* fp = A.cpl + size_t(2 * c) * A.pw;
    const uint2* rp = fp + A.pw;
    const uint2* p0 = (pos ? A.ap : fp) + off;  // shifted side, forward candidate
    const uint2* p1 = (pos ? A.ap : rp) + off;  // shifted side, reverse-complement candidate
    const uint2* q0 = (pos ? fp : A.ap) + 1;    // word-aligned side
    const uint2* q1 = (pos ? rp : A.ap) + 1;
    const u32 lastMask = (ov & 31) ? ~(~0u >> (ov & 31)) : ~0u;
    // one word per iteration: the lanes are mostly full here, and fewer
    // words in flight keep the registers of the calling kernels
    for (int j = 2; __any(j < nw && (mf < lim || mr < lim)); j++) {
        const uint2 a0 = p0[j], a1 = p0[j + 1], b0 = p1[j], b1 = p1[j + 1];
        const uint2 va = q0[j], vb = q1[j];
        const u32 msk = j + 1 < nw ? ~0u : (j < nw ? lastMask : 0u);
        const u32 xf = (__builtin_amdgcn_alignbit(a0.x, a1.x, sh) ^ va.x) |
                       (__builtin_amdgcn_alignbit(a0.y, a1.y, sh) ^ va.y);
        const u32 xr = (__builtin_amdgcn_alignbit(b0.x, b1.x, sh) ^ vb.x) |
                       (__builtin_amdgcn_alignbit(b0.y, b1.y, sh) ^ vb.y);
        mf += __popc(xf & msk);
        mr += __popc(xr & msk);
    }
    if (valid) {
        const u32 base = u32(la + lc - 2 * ov);
        if (mf < lim) atomicMin(&A.keys[2 * c], ((u32(mf) + base) << 11) | u32(u));
        if (mr < lim) atomicMin(&A.keys[2 * c + 1], ((u32(mr) + base) << 11) | u32(u));
    }
    wave_sync();
}

// appends the lanes in `survive` (their order index u and mismatch counts)
// to the wave's survivor slice, finishing the queued ones first when the
// slice would overflow; `count` is wave-uniform
__device__ __forceinline__ void queue_survivors(const AlignCtx& A, int c, bool survive, int u, int mf, int mr,
                                                int& count) {
    const u64 bal = __ballot(survive);
    if (!bal) return;
    const int k = __popcll(bal);
    if (count + k > kSurvivors) {
        finish_survivors(A, count);
        count = 0;
    }
    const int lane = lane_id();
    if (survive) {
        const int slot = count + __popcll(bal & (lane == 0 ? 0ull : (~0ull >> (64 - lane))));
        A.survA[slot] = (u32(c) << 11) | u32(u);
        A.survB[slot] = u32(mf) | (u32(mr) << 16);
    }
    count += k;
    wave_sync();
}

// Alignment of the anchor against the candidate planes of one wave, all
// shifts in parallel, both orientations of a shift in the same lane; key =
// score << 11 | order index u (u = s for the shifts s >= 0, then
// u = nPos - 1 - s for s < 0), combined into keys[2c], keys[2c + 1] by
// atomic minimum (both set to ~0u by the caller first).  The shifts s >= 0
// and s < 0 run in separate wave-uniform loops.  For s >= 0 the anchor is
// read from base s against the candidate from base 0, so the shifted anchor
// words serve both orientations; for s < 0 the candidate is read from base
// -s against the anchor from base 0, whose words are the same in every
// lane.  Every shift compares its first 64 bases here; the few shifts still
// under their mismatch limit (in either orientation) with more bases to go
// are queued and finished later a wavefront at a time (finish_survivors),
// instead of holding a wavefront of mostly finished lanes.  Mismatch counts
// only grow, so counting past the limit changes no result; reads past a
// lane's last word stay inside the workgroup's LDS and are masked.
__device__ __forceinline__ void align_wave(const AlignCtx& A, int c, int lc, int& count) {
    const int lane = lane_id();
    const int la = A.la;
    const int nPos = max(0, la - A.minov + 1);
    const int nNeg = max(0, lc - A.minov);
    const int total = la + lc;
    const uint2* ap = A.ap;
    const uint2* fp = A.cpl + size_t(2 * c) * A.pw;
    const uint2* rp = fp + A.pw;
    u32 kf = ~0u, kr = ~0u;
    for (int u0 = 0; u0 < nPos; u0 += kWave) {
        const int s = u0 + lane;  // order index u = s
        const int ov = min(la - s, lc);
        const int lim = int(float(ov) * A.maxErrorRate);
        const int nw = (s < nPos && lim > 0) ? (ov + 31) >> 5 : 0;
        const u32 sh = u32(32 - (s & 31)) & 31u;
        const uint2* p = ap + (s >> 5) + ((s & 31) != 0);
        const u32 lastMask = (ov & 31) ? ~(~0u >> (ov & 31)) : ~0u;
        int mf = 0, mr = 0;
        // the first two words: one LDS round trip covers 64 bases
        if (__any(nw > 0)) {
            const uint2 a0 = p[0], a1 = p[1], a2 = p[2];
            const uint2 f1 = fp[1], f2 = fp[2], r1 = rp[1], r2 = rp[2];
            u32 msk, msk2;
            word_masks(0, nw, lastMask, &msk, &msk2);
            const u32 h1 = __builtin_amdgcn_alignbit(a0.x, a1.x, sh), l1 = __builtin_amdgcn_alignbit(a0.y, a1.y, sh);
            const u32 h2 = __builtin_amdgcn_alignbit(a1.x, a2.x, sh), l2 = __builtin_amdgcn_alignbit(a1.y, a2.y, sh);
            mf += __popc(((h1 ^ f1.x) | (l1 ^ f1.y)) & msk) + __popc(((h2 ^ f2.x) | (l2 ^ f2.y)) & msk2);
            mr += __popc(((h1 ^ r1.x) | (l1 ^ r1.y)) & msk) + __popc(((h2 ^ r2.x) | (l2 ^ r2.y)) & msk2);
        }
        const bool survive = nw > 2 && (mf < lim || mr < lim);
        // finished under the limit (rare: most shifts end over it)
        if (nw > 0 && !survive && (mf < lim || mr < lim)) {
            const u32 base = u32(total - 2 * ov);
            if (mf < lim) kf = min(kf, ((u32(mf) + base) << 11) | u32(s));
            if (mr < lim) kr = min(kr, ((u32(mr) + base) << 11) | u32(s));
        }
        queue_survivors(A, c, survive, s, mf, mr, count);
    }
    for (int v0 = 0; v0 < nNeg; v0 += kWave) {
        const int d = v0 + lane + 1;  // shift s = -d, order index u = nPos - 1 + d
        const int ov = min(la, lc - d);
        const int lim = int(float(ov) * A.maxErrorRate);
        const int nw = (d <= nNeg && lim > 0) ? (ov + 31) >> 5 : 0;
        const u32 sh = u32(32 - (d & 31)) & 31u;
        const int off = (d >> 5) + ((d & 31) != 0);
        const uint2* pf = fp + off;
        const uint2* pr = rp + off;
        const u32 lastMask = (ov & 31) ? ~(~0u >> (ov & 31)) : ~0u;
        int mf = 0, mr = 0;
        if (__any(nw > 0)) {
            const uint2 f0 = pf[0], f1 = pf[1], f2 = pf[2], r0 = pr[0], r1 = pr[1], r2 = pr[2];
            const uint2 a1 = ap[1], a2 = ap[2];
            u32 msk, msk2;
            word_masks(0, nw, lastMask, &msk, &msk2);
            const u32 xf = (__builtin_amdgcn_alignbit(f0.x, f1.x, sh) ^ a1.x) |
                           (__builtin_amdgcn_alignbit(f0.y, f1.y, sh) ^ a1.y);
            const u32 xr = (__builtin_amdgcn_alignbit(r0.x, r1.x, sh) ^ a1.x) |
                           (__builtin_amdgcn_alignbit(r0.y, r1.y, sh) ^ a1.y);
            const u32 xf2 = (__builtin_amdgcn_alignbit(f1.x, f2.x, sh) ^ a2.x) |
                            (__builtin_amdgcn_alignbit(f1.y, f2.y, sh) ^ a2.y);
            const u32 xr2 = (__builtin_amdgcn_alignbit(r1.x, r2.x, sh) ^ a2.x) |
                            (__builtin_amdgcn_alignbit(r1.y, r2.y, sh) ^ a2.y);
            mf += __popc(xf & msk) + __popc(xf2 & msk2);
            mr += __popc(xr & msk) + __popc(xr2 & msk2);
        }
        const int u = nPos - 1 + d;
        const bool survive = nw > 2 && (mf < lim || mr < lim);
        if (nw > 0 && !survive && (mf < lim || mr < lim)) {
            const u32 base = u32(total - 2 * ov);
            if (mf < lim) kf = min(kf, ((u32(mf) + base) << 11) | u32(u));
            if (mr < lim) kr = min(kr, ((u32(mr) + base) << 11) | u32(u));
        }
        queue_survivors(A, c, survive, u, mf, mr, count);
    }
    // few lanes hold a key here, so each adds its own instead of a
    // wavefront reduction per candidate
    if (kf != ~0u) atomicMin(&A.keys[2 * c], kf);
    if (kr != ~0u) atomicMin(&A.keys[2 * c + 1], kr);
    wave_sync();
}

struct Aln {
    int overlap, shift, nOps;
    bool valid;
};

__device__ __forceinline__ Aln decode_key(u32 key, int la, int lc, int minov) {
    Aln a;
    const int total = la + lc;
    if (key == ~0u) {
        a.valid = false;
        a.shift = -lc;
        a.overlap = 0;
        a.nOps = 0;
        return a;
    }
    const int nPos = max(0, la - minov + 1);
    const int u = int(key & 2047u);
    const int score = int(key >> 11);
    a.shift = u < nPos ? u : -(u - nPos + 1);
    const int cb = max(-a.shift, 0);
    const int ce = min(lc, la - a.shift);
    a.overlap = ce - cb;
    a.nOps = score - total + 2 * a.overlap;
    a.valid = true;
    return a;
}

// include/alignmentorientation.hpp:14-70; maxErr = 4 * estimatedErrorrate
__device__ __forceinline__ int choose_orientation(const Aln& f, const Aln& r, int la, float minOverlapRatio,
                                                  int minOverlap, float maxErr) {
    const int prod = int(float(la) * minOverlapRatio);
    const int minov = prod > minOverlap ? prod : minOverlap;
    if (f.valid && f.overlap >= minov) {
        if (r.valid && r.overlap >= minov) {
            const float ratio = float(f.nOps) / float(f.overlap);
            const float rratio = float(r.nOps) / float(r.overlap);
            if (ratio < rratio) return ratio < maxErr ? 1 : 3;
            if (rratio < ratio) return rratio < maxErr ? 2 : 3;
            if (ratio < maxErr) return f.overlap > r.overlap ? 1 : 2;
            return 3;
        }
        return float(f.nOps) / float(f.overlap) < maxErr ? 1 : 3;
    }
    if (r.valid && r.overlap >= minov && float(r.nOps) / float(r.overlap) < maxErr) return 2;
    return 3;
}

// include/forest.hpp:57-68, 211-226: mean leaf probability over the trees,
// accepting as soon as the running mean reaches the threshold
// candidate record from its staged id / length / pair flag and the two
// alignment keys: orientation chosen, shift, overlap and mismatches of it
__device__ __forceinline__ CandMeta aligned_meta(const CandMeta& st, u32 kf, u32 kr, int la, int minov,
                                                 const CorrectParams& P) {
    const int lc = st.len;
    const Aln f = decode_key(kf, la, lc, minov);
    const Aln r = decode_key(kr, la, lc, minov);
    const int flag = choose_orientation(f, r, la, P.minOverlapRatio, P.minOverlap, P.maxErr4);
    const Aln& ch = flag == 2 ? r : f;
    CandMeta m;
    m.id = st.id;
    m.shift = short(ch.shift);
    m.len = u16(lc);
    m.u.al.nOps = u16(ch.nOps);
    m.u.al.overlap = u16(ch.overlap);
    m.flag = u8(flag);
    m.paired = st.paired;
    m.pad = 0;
    return m;
}

__device__ bool forest_decide(const DevForest& F, const float* f) {
    float sum = 0.f;
    for (int t = 0; t < F.numTrees; t++) {
        const ForestNode* tree = F.nodes + F.treeStart[t];
        u32 i = 0;
        float p;
        for (;;) {
            const ForestNode nd = tree[i];
            if (double(f[nd.att]) <= nd.split) {
                if (nd.flag & 2) {
                    p = __uint_as_float(nd.lhs);
                    break;
                }
                i = nd.lhs;
            } else {
                if (nd.flag & 1) {
                    p = __uint_as_float(nd.rhs);
                    break;
                }
                i = nd.rhs;
            }
        }
        sum += p;
        if (sum / float(F.numTrees) >= F.threshold) return true;
    }
    return false;
}

struct ColumnStats {
    float avgSupport, minSupport;
    int minCov, maxCov;
};

// MSA properties over columns [first, last) (src/msa.cpp:304-331) by one
// wavefront (all lanes call): minima and maxima over the lanes (support is
// non-negative, so its bit pattern orders like the value), the support sum
// added in column order by lane 0 alone (sequential FP32, as on the CPU);
// avgSupport is valid in lane 0, the other fields in every lane
__device__ ColumnStats column_stats_wave(const Lds& L, int first, int last) {
    const int lane = lane_id();
    u32 mnBits = __float_as_uint(L.support[first]);
    int mnc = L.coverage[first], mxc = mnc;
    for (int c = first + lane; c < last; c += kWave) {
        const u32 s = __float_as_uint(L.support[c]);
        mnBits = s < mnBits ? s : mnBits;
        const int cv = L.coverage[c];
        mnc = cv < mnc ? cv : mnc;
        mxc = cv > mxc ? cv : mxc;
    }
    ColumnStats cs;
    cs.minSupport = __uint_as_float(wave_min_u32(mnBits));
    cs.minCov = wave_min_int(mnc);
    cs.maxCov = wave_max_int(mxc);
    float sum = 0.0f;
    if (lane == 0) {
#pragma unroll 8
        for (int c = first; c < last; c++) sum += L.support[c];
    }
    cs.avgSupport = sum / float(last - first);
    return cs;
}

// extract_anchor_v3 / extract_cands_v3 (include/classification.hpp:748-877);
// origBase is the anchor's (or candidate's) base at column pos
__device__ void column_features(const Lds& L, int cols, int pos, u32 origBase, const ColumnStats& cs, float cov,
                                float* f) {
    const float wA = L.wgt[0 * cols + pos], wC = L.wgt[1 * cols + pos], wG = L.wgt[2 * cols + pos],
                wT = L.wgt[3 * cols + pos];
    const float wsum = wA + wC + wG + wT;
    const u32 cons = L.cons[pos];
    const float wCons = L.wgt[cons * cols + pos];
    const float cCons = float(L.cnt[cons * cols + pos]);
    const float wOrig = L.wgt[origBase * cols + pos];
    const float cOrig = float(L.cnt[origBase * cols + pos]);
    const float coverage = float(L.coverage[pos]);
    f[0] = cOrig / coverage;
    f[1] = wOrig / wsum;
    f[2] = wOrig / cOrig;
    f[3] = cCons / coverage;
    f[4] = L.support[pos];
    f[5] = wCons / cCons;
    f[6] = wsum / coverage;
    f[7] = coverage / cov;
    f[8] = wsum / cov;
    f[9] = cs.avgSupport;
    f[10] = cs.minSupport;
    f[11] = float(cs.maxCov) / cov;
    f[12] = float(cs.minCov) / cov;
}

// The 32 gate features of position i of a read in the MSA (column ab + i)
// whose consensus differs from the read's base: the anchor, or a candidate
// in its MSA orientation.  base(ii) is the read's base at position ii, qwi
// the quality weight of its base i (1 without qualities), `acs` the MSA
// properties over the read's columns, nd the read's number of decision
// columns, m the candidates in the MSA.  Shared with the oracle's
// restatement.
template <class BaseAt>
__device__ void gate_features(const Lds& L, int cols, int ab, int la, int i, int nd, int m, const ColumnStats& acs,
                              float estCov, float qwi, const BaseAt& base, float* f) {
    const int col = ab + i;
    const u32 o = base(i);
    const u32 c = L.cons[col];
    const float cov = float(L.coverage[col]);
    const float w0 = L.wgt[0 * cols + col], w1 = L.wgt[1 * cols + col], w2 = L.wgt[2 * cols + col],
                w3 = L.wgt[3 * cols + col];
    const float n0 = float(L.cnt[0 * cols + col]), n1 = float(L.cnt[1 * cols + col]), n2 = float(L.cnt[2 * cols + col]),
                n3 = float(L.cnt[3 * cols + col]);
    const float wsum = w0 + w1 + w2 + w3;
    const float wo = o == 0 ? w0 : o == 1 ? w1 : o == 2 ? w2 : w3;
    const float wc = c == 0 ? w0 : c == 1 ? w1 : c == 2 ? w2 : w3;
    const float no = o == 0 ? n0 : o == 1 ? n1 : o == 2 ? n2 : n3;
    const float nc = c == 0 ? n0 : c == 1 ? n1 : c == 2 ? n2 : n3;
    f[0] = no / cov;
    f[1] = wo / wsum;
    f[2] = nc / cov;
    f[3] = L.support[col];
    f[4] = wo / fmaxf(no, 1.0f);
    f[5] = wc / fmaxf(nc, 1.0f);
    f[6] = cov / estCov;
    f[7] = wsum / estCov;
    f[8] = (cov - no - nc) / cov;
    f[9] = qwi;
    f[10] = float(i) / float(la);
    f[11] = float(min(i, la - 1 - i)) / float(la);
    f[12] = acs.avgSupport;
    f[13] = acs.minSupport;
    f[14] = float(acs.minCov) / estCov;
    f[15] = float(acs.maxCov) / estCov;
    f[16] = float(m) / estCov;
    const int dcol[4] = {-2, -1, 1, 2};
    for (int k = 0; k < 4; k++) {
        const int cc = col + dcol[k];
        f[17 + k] = (cc >= 0 && cc < cols) ? L.support[cc] : 0.0f;
        const int ii = i + dcol[k];
        f[21 + k] = (ii >= 0 && ii < la && u32(L.cons[ab + ii]) != base(ii)) ? 1.0f : 0.0f;
    }
    f[25] = float(nd) / float(la);
    f[26] = col - 1 >= 0 ? float(L.coverage[col - 1]) / estCov : 0.0f;
    f[27] = col + 1 < cols ? float(L.coverage[col + 1]) / estCov : 0.0f;
    f[28] = (wc - wo) / wsum;
    const float q0 = n0 / cov, q1 = n1 / cov, q2 = n2 / cov, q3 = n3 / cov;
    f[29] = q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3;
    f[30] = wsum / cov;
    int near = 0;
    for (int d = -5; d <= 5; d++) {
        const int ii = i + d;
        if (d != 0 && ii >= 0 && ii < la && u32(L.cons[ab + ii]) != base(ii)) near++;
    }
    f[31] = float(near) / 10.0f;
}

// Per-wave LDS of the candidate gate: feature rows f32 [16][33], input tile
// bf16 [16][32], logits [16], MLP scratch, then one accept flag per position.
constexpr int kCandGateTile = 16 * (kMlpIn + 1) * 4 + 16 * kMlpIn * 2 + 16 * 4 + kMlpTileScratch;
__host__ __device__ inline int cand_gate_scratch(int maxLen) { return kCandGateTile + (maxLen + 15) / 16 * 16; }

// Learned gate over the decision columns of one candidate (positions where
// the consensus differs from its base), 16 at a time in position order, by
// one wavefront; acc[i] = 1 if the gate accepts the consensus at i.  Decision
// columns are recorded when the engine records candidate columns.
__device__ void gate_candidate(const CorrectParams& P, const Lds& L, const DevReadStore& rs, int cols, int qb,
                               const CandMeta& cm, int m, const ColumnStats& ccs, char* ws, u8* acc) {
    const int lane = lane_id();
    const int len = cm.len;
    const bool rc = cm.flag == 2;
    const u32* enc = rs.seqs + u64(cm.id) * rs.pitch;
    const u32* qrow = P.useQuality ? rs.quals + u64(cm.id) * rs.qpitch : nullptr;
    auto base = [&](int ii) { return oriented_base(enc, len, ii, rc); };
    float* ft = reinterpret_cast<float*>(ws);                                      // [16][33]
    bf16raw* xt = reinterpret_cast<bf16raw*>(ws + 16 * (kMlpIn + 1) * 4);          // [16][32]
    float* zt = reinterpret_cast<float*>(ws + 16 * (kMlpIn + 1) * 4 + 16 * kMlpIn * 2);  // [16]
    char* scr = ws + 16 * (kMlpIn + 1) * 4 + 16 * kMlpIn * 2 + 16 * 4;
    const u64 below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int nd = 0;
    for (int i0 = 0; i0 < len; i0 += kWave) {
        const int i = i0 + lane;
        const bool d = i < len && u32(L.cons[qb + i]) != base(i);
        nd += __popcll(__ballot(d));
        if (i < len) acc[i] = 0;
    }
    wave_sync();
    MlpWeights W;
    W.w1 = P.gateCW1;
    W.w2 = P.gateCW2;
    W.w3 = P.gateCW3;
    if (P.gateCParams) {
        W.b1 = P.gateCParams + kMlpB1;
        W.b2 = P.gateCParams + kMlpB2;
        W.b3 = P.gateCParams + kMlpB3;
        W.w4 = P.gateCParams + kMlpW4;
        W.b4 = P.gateCParams[kMlpB4];
    }
    for (int i0 = 0; i0 < len; i0 += kWave) {
        const int i = i0 + lane;
        const bool d = i < len && u32(L.cons[qb + i]) != base(i);
        u64 bal = __ballot(d);
        while (bal) {  // uniform: bal is the same in every lane
            const int ng = min(16, __popcll(bal));
            const int rank = __popcll(bal & below);
            const bool mine = ((bal >> lane) & 1ull) && rank < 16;
            if (mine) {
                float* f = ft + rank * (kMlpIn + 1);
                const float qwi =
                    qrow ? L.qw[(unsigned char)quality_char(qrow, rc ? rs.maxLen - 1 - i : i, rs.qbits)] : 1.0f;
                gate_features(L, cols, qb, len, i, nd, m, ccs, P.estimatedCoverage, qwi, base, f);
                for (int k = 0; k < kMlpIn; k++) xt[rank * kMlpIn + k] = f2bf(f[k]);
                if (P.gateFeat && P.gateRecordCands) {
                    const int slot = atomicAdd(P.gateCount, 1);
                    if (slot < P.gateCap) {
                        // forward-strand position and bases of the candidate read
                        const u32 o = base(i), cn = L.cons[qb + i];
                        for (int k = 0; k < kMlpIn; k++) P.gateFeat[u64(slot) * kMlpIn + k] = f[k];
                        P.gateIds[slot] = cm.id;
                        P.gatePos[slot] = u16(rc ? len - 1 - i : i);
                        P.gateOrig[slot] = u8(rc ? 3u - o : o);
                        P.gateCons[slot] = u8(rc ? 3u - cn : cn);
                    }
                }
            }
            if (lane < 16 && lane >= ng)
                for (int k = 0; k < kMlpIn; k++) xt[lane * kMlpIn + k] = 0;
            wave_sync();
            if (P.gateCParams) mlp_tile_forward(W, xt, scr, zt);
            wave_sync();
            if (mine) acc[i] = (P.gateCParams && zt[rank] >= P.gateCLogit) ? 1 : 0;
            wave_sync();
            for (int k = 0; k < ng; k++) bal &= bal - 1;
        }
    }
    wave_sync();
}

}  // namespace

// kGated: the learned-gate mode (correctionType 3), kGateCands: with the
// gate on candidate columns too (correctionTypeCands 3), kRefine: refinement
// rounds (the plain modes; in the gated mode only to record the decision
// columns of the plain method); compiled separately so that their register
// use does not lower the occupancy of the other modes.
template <bool kGated, bool kGateCands, bool kRefine = !kGated>
__global__ void __launch_bounds__(256) k_correct(CorrectParams P, LdsDims dims) {
    HIP_DYNAMIC_SHARED(uint4, ldsRaw)  // uint4: 16-byte aligned dynamic LDS
    const Lds L = carve(reinterpret_cast<char*>(ldsRaw), dims);
    __shared__ int sAnchor, sMin, sMax, sCount[3], sFlag, sCol, sTotal;
    __shared__ float sThr;
    const DevReadStore& rs = P.rs;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid / kWave;
    const int nwaves = blockDim.x / kWave;
    const int pw = dims.pw();

    if (P.useQuality) {
        for (int i = tid; i < 256; i += blockDim.x) L.qw[i] = P.qweights[i];
    }
    for (;;) {
        if (tid == 0) sAnchor = atomicAdd(P.queueHead, 1);
        __syncthreads();
        const int a = sAnchor;
        __syncthreads();
        if (a >= P.numAnchors) break;

        const u64 id = P.firstAnchor + u64(a);
        CARE_CHECK(id < rs.n);
        const int la = rs.lens[id];
        const int n = P.candCount[a];
        CARE_CHECK(la >= 1 && la <= dims.maxLen && n >= 0 && n <= P.maxCand && rs.pitch == dims.pitch);
        const u32* ids = P.candIds + u64(a) * P.maxCand;
        u8 status = 0;
        char* outSeq = P.outSeq + u64(a) * P.outPitch;

        if (n == 0) {
            if (tid == 0) P.outStatus[a] = 0;
            continue;
        }

        // anchor words, planes and qualities
        const u32* aenc = rs.seqs + id * rs.pitch;
        for (int i = tid; i < rs.pitch; i += blockDim.x) L.aw[i] = aenc[i];
        if (P.useQuality) {
            const u32* q = rs.quals + id * rs.qpitch;
            for (int i = tid; i < la; i += blockDim.x) L.anchorQ[i] = quality_char(q, i, rs.qbits);
        }
        __syncthreads();
        for (int p = tid; p < pw; p += blockDim.x) {
            const int w = p - 1;  // plane word (index 0 is padding)
            u32 hi = 0, lo = 0;
            if (w >= 0 && 2 * w < rs.pitch) {
                const u32 x0 = L.aw[2 * w];
                const u32 x1 = 2 * w + 1 < rs.pitch ? L.aw[2 * w + 1] : 0u;
                hi = (even_bits(x0 >> 1) << 16) | even_bits(x1 >> 1);
                lo = (even_bits(x0) << 16) | even_bits(x1);
                // clear bits beyond the read
                const int valid = la - 32 * w;
                if (valid <= 0) { hi = 0; lo = 0; }
                else if (valid < 32) {
                    const u32 mask = ~0u << (32 - valid);
                    hi &= mask;
                    lo &= mask;
                }
            }
            L.ap[p] = make_uint2(hi, lo);
        }
        __syncthreads();

        // ---- phase A: alignments, one wave per candidate ----
        const int minovA = max(P.minOverlap, int(float(la) * P.minOverlapRatio));
        const u32* mateIds = nullptr;
        int mateN = 0;
        if (P.pairedEnd) {
            const int mate = a ^ 1;
            if (mate < P.numAnchors) {
                mateIds = P.candIds + u64(mate) * P.maxCand;
                mateN = P.candCount[mate];
            }
        }
        CandMeta* metaA = P.metaA + u64(a) * P.maxCand;
        CandMeta* metaB = P.metaB + u64(a) * P.maxCand;
        // candidate ids, lengths, 2-bit rows and pair flags of a chunk are
        // gathered into LDS by the whole block, so their global loads (and
        // the binary searches in the mate's candidate list) run in parallel
        // the mate's candidate list is searched in LDS when it fits the
        // alignment keys and survivor slices, which are free until this
        // chunk's alignment starts (not in the gated modes, where its
        // registers would cost a wavefront per SIMD)
        const int mateCap = kGated ? -1 : 2 * dims.chunk + 2 * kSurvivors * dims.waves;
        for (int c0 = 0; c0 < n; c0 += dims.chunk) {
            const int nc = min(dims.chunk, n - c0);
            __syncthreads();  // the previous chunk's rows and keys are no longer read
            if (mateIds && mateN <= mateCap) {
                for (int i = tid; i < mateN; i += blockDim.x) L.keys[i] = mateIds[i];
                __syncthreads();
            }
            for (int i = tid; i < nc * rs.pitch; i += blockDim.x) {
                const int c = i / rs.pitch, w = i - c * rs.pitch;
                const u32 cid = ids[c0 + c];
                CARE_CHECK(cid < rs.n && cid != u32(id) && c < dims.chunk);
                L.rows[i] = rs.seqs[u64(cid) * rs.pitch + w];
                if (w == 0) {
                    L.mst[c].id = cid;
                    L.mst[c].len = u16(rs.lens[cid]);
                    // the candidate's mate is a candidate of the anchor's mate
                    // (include/corrector.hpp:763-807)
                    u8 paired = 0;
                    if (mateIds) {
                        const u32 want = cid ^ 1u;
                        CARE_CHECK(mateN >= 0 && mateN <= P.maxCand);
                        int lo2 = 0, hi2 = mateN;
                        if (mateN <= mateCap) {
                            while (lo2 < hi2) {
                                const int mid = (lo2 + hi2) >> 1;
                                if (L.keys[mid] < want) lo2 = mid + 1;
                                else hi2 = mid;
                            }
                            paired = (lo2 < mateN && L.keys[lo2] == want) ? 1 : 0;
                        } else {
                            while (lo2 < hi2) {
                                const int mid = (lo2 + hi2) >> 1;
                                if (mateIds[mid] < want) lo2 = mid + 1;
                                else hi2 = mid;
                            }
                            paired = (lo2 < mateN && mateIds[lo2] == want) ? 1 : 0;
                        }
                    }
                    L.mst[c].paired = paired;
                }
            }
            __syncthreads();
            // bit planes of every staged candidate: forward ...
            for (int it = tid; it < nc * pw; it += blockDim.x) {
                const int c = it / pw, p = it - c * pw;
                const int lc = L.mst[c].len;
                const u32* cenc = L.rows + c * rs.pitch;
                const int w = p - 1;  // plane word (index 0 is padding)
                u32 hi = 0, lo = 0;
                if (w >= 0 && w < (lc + 31) / 32) {
                    const u32 x0 = cenc[2 * w];
                    const u32 x1 = 2 * w + 1 < rs.pitch ? cenc[2 * w + 1] : 0u;
                    hi = (even_bits(x0 >> 1) << 16) | even_bits(x1 >> 1);
                    lo = (even_bits(x0) << 16) | even_bits(x1);
                    const int valid = lc - 32 * w;
                    if (valid < 32) {
                        const u32 mask = ~0u << (32 - valid);
                        hi &= mask;
                        lo &= mask;
                    }
                }
                L.cpl[size_t(2 * c) * pw + p] = make_uint2(hi, lo);
            }
            __syncthreads();
            // ... and reverse complement: rc[q] = 1 - f[lc-1-q]
            for (int it = tid; it < nc * pw; it += blockDim.x) {
                const int c = it / pw, p = it - c * pw;
                const int lc = L.mst[c].len;
                const uint2* fp = L.cpl + size_t(2 * c) * pw;
                const int w = p - 1;
                u32 hi = 0, lo = 0;
                if (w >= 0 && w < (lc + 31) / 32) {
                    const int e = lc - 32 * (w + 1);  // first fwd bit of the window
                    u32 xh, xl;
                    if (e >= 0) {
                        const int k = e >> 5, r = e & 31;
                        const uint2 a0 = fp[1 + k], a1 = fp[2 + k];
                        xh = r ? (a0.x << r) | (a1.x >> (32 - r)) : a0.x;
                        xl = r ? (a0.y << r) | (a1.y >> (32 - r)) : a0.y;
                    } else {
                        xh = fp[1].x >> (-e);
                        xl = fp[1].y >> (-e);
                    }
                    hi = ~__brev(xh);
                    lo = ~__brev(xl);
                    const int valid = lc - 32 * w;
                    if (valid < 32) {
                        const u32 mask = ~0u << (32 - valid);
                        hi &= mask;
                        lo &= mask;
                    }
                }
                L.cpl[size_t(2 * c + 1) * pw + p] = make_uint2(hi, lo);
            }
            __syncthreads();
            // each wave aligns a candidate at a time and leaves its two keys
            // in LDS; the keys of the whole chunk are then decoded and the
            // orientations chosen one thread per candidate (decoding in the
            // aligning wave's lane 0 would issue every instruction for the
            // whole wave)
            {
                AlignCtx A;
                A.ap = L.ap;
                A.cpl = L.cpl;
                A.mst = L.mst;
                A.keys = L.keys;
                A.survA = L.keys + 2 * dims.chunk + 2 * kSurvivors * wave;
                A.survB = A.survA + kSurvivors;
                A.la = la;
                A.pw = pw;
                A.minov = minovA;
                A.maxErrorRate = P.maxErrorRate;
                int count = 0;  // queued survivors of this wave (wave-uniform)
                for (int c = wave; c < nc; c += nwaves) {
                    if (lane == 0) {
                        CARE_CHECK(c0 + c < n);
                        L.keys[2 * c] = ~0u;
                        L.keys[2 * c + 1] = ~0u;
                    }
                    wave_sync();
                    align_wave(A, c, L.mst[c].len, count);
                }
                if (count > 0) finish_survivors(A, count);
            }
            __syncthreads();
            for (int c = tid; c < nc; c += blockDim.x)
                metaA[c0 + c] = aligned_meta(L.mst[c], L.keys[2 * c], L.keys[2 * c + 1], la, minovA, P);
        }
        __syncthreads();

        // ---- phase B: orientation + mismatch-ratio filters ----
        if (!P.pairedEnd) {
            if (tid < 3) sCount[tid] = 0;
            __syncthreads();
            int c0 = 0, c1 = 0, c2 = 0;
            for (int c = tid; c < n; c += blockDim.x) {
                const CandMeta m = metaA[c];
                if (m.flag == 3) continue;
                const float ratio = float(m.u.al.nOps) / float(m.u.al.overlap);
                c0 += ratio < P.mmBase2;
                c1 += ratio < P.mmBase3;
                c2 += ratio < P.mmBase4;
            }
            c0 = wave_sum_int(c0);
            c1 = wave_sum_int(c1);
            c2 = wave_sum_int(c2);
            if (lane == 0) {
                atomicAdd(&sCount[0], c0);
                atomicAdd(&sCount[1], c1);
                atomicAdd(&sCount[2], c2);
            }
            __syncthreads();
            if (tid == 0) {
                float thr = -3.40282347e38f;  // numeric_limits<float>::lowest()
                if (sCount[0] > 0 || sCount[1] > 0 || sCount[2] > 0) {
                    if (float(sCount[0]) >= P.goodAlignmentsThreshold) thr = P.mmBase2;
                    else if (float(sCount[1]) >= P.goodAlignmentsThreshold) thr = P.mmBase3;
                    else if (float(sCount[2]) >= P.goodAlignmentsThreshold) thr = P.mmBase4;
                }
                sThr = thr;
            }
            __syncthreads();
        }
        int m = 0;
        for (int base = 0; base < n; base += blockDim.x) {
            const int c = base + tid;
            bool keep = false;
            CandMeta cm{};
            if (c < n) {
                cm = load_meta(metaA + c);
                if (cm.flag != 3) {
                    const float ratio = float(cm.u.al.nOps) / float(cm.u.al.overlap);
                    keep = P.pairedEnd ? (cm.paired || ratio < P.pairedThreshold) : (ratio < sThr);
                }
            }
            int cnt;
            const int pos = block_exclusive_count(keep, L.scratch, &cnt);
            if (keep) {
                CARE_CHECK(m + pos < n && cm.u.al.overlap > 0);
                const float w = 1.0f - sqrtf(float(cm.u.al.nOps) / (float(cm.u.al.overlap) * P.maxErrorRate));
                cm.u.weight = w;
                store_meta(metaB + m + pos, cm);
            }
            m += cnt;
        }
        __syncthreads();
        if (m == 0) {
            if (tid == 0) P.outStatus[a] = 0;
            continue;
        }

        // ---- phase C: MSA + refinement (include/corrector.hpp:1307-1421) ----
        CandMeta* cur = metaB;
        CandMeta* other = metaA;
        // MSA columns read beyond the anchor: candidate correction reaches
        // new_columns past it, gate features two columns past a corrected read
        const int margin = P.correctCandidates ? P.newColumns + 2 : 2;
        Msa msa = build_msa<!kGated>(P, L, dims, cur, m, la, margin, &sMin, &sMax);
        const int maxIter = kRefine ? P.maxRefinements : 0;  // the gated mode skips refinement
        for (int iter = 0; iter < maxIter && m > 0; iter++) {
            // src/msa.cpp:532-734
            if (tid == 0) {
                sFlag = 0;
                sCol = 0x7FFFFFFF;
            }
            __syncthreads();
            for (int i = tid; i < la; i += blockDim.x)
                if (L.cons[msa.ab + i] != u8(base_at(L.aw, i))) sFlag = 1;
            __syncthreads();
            if (!sFlag) break;
            int firstCol = 0x7FFFFFFF;
            for (int i = tid; i < la; i += blockDim.x) {
                const int col = msa.ab + i;
                const int ci = L.cons[col];
                bool found = false;
                for (int b = 0; b < 4; b++)
                    if (b != ci && P.sigCountThreshold <= int(L.cnt[b * msa.cols + col])) found = true;
                if (found) firstCol = min(firstCol, col);
            }
            firstCol = wave_min_int(firstCol);
            if (lane == 0 && firstCol != 0x7FFFFFFF) atomicMin(&sCol, firstCol);
            __syncthreads();
            const int col = sCol;
            if (col == 0x7FFFFFFF) break;
            const int ci = L.cons[col];
            int foundIdx = -1;
            for (int b = 0; b < 4; b++)
                if (b != ci && P.sigCountThreshold <= int(L.cnt[b * msa.cols + col])) foundIdx = b;
            const u32 origBase = base_at(L.aw, col - msa.ab);
            const bool keepMatching = origBase == u32(foundIdx);
            if (tid == 0) sFlag = 0;  // very good alignment among removed
            __syncthreads();
            int removed = 0;
            for (int c = tid; c < m; c += blockDim.x) {
                const CandMeta cm = cur[c];
                const int rb = msa.ab + cm.shift;
                const int re = rb + cm.len;
                bool rem = false;
                if (col >= rb && col < re) {
                    const u32 b = oriented_base(rs.seqs + u64(cm.id) * rs.pitch, cm.len, col - rb, cm.flag == 2);
                    rem = keepMatching ? (b != u32(foundIdx)) : (b == u32(foundIdx));
                }
                if (rem && cm.u.weight >= 0.90f) sFlag = 1;
                removed += rem;
            }
            removed = wave_sum_int(removed);
            if (lane == 0 && removed) atomicAdd(&sFlag, 2);
            __syncthreads();
            if ((sFlag & 1) || sFlag == 0) break;  // nothing removed: MSA unchanged
            int m2 = 0;
            for (int base = 0; base < m; base += blockDim.x) {
                const int c = base + tid;
                bool keep = false;
                CandMeta cm{};
                if (c < m) {
                    cm = load_meta(cur + c);
                    const int rb = msa.ab + cm.shift;
                    const int re = rb + cm.len;
                    keep = true;
                    if (col >= rb && col < re) {
                        const u32 b = oriented_base(rs.seqs + u64(cm.id) * rs.pitch, cm.len, col - rb, cm.flag == 2);
                        keep = !(keepMatching ? (b != u32(foundIdx)) : (b == u32(foundIdx)));
                    }
                }
                int cnt;
                const int pos = block_exclusive_count(keep, L.scratch, &cnt);
                CARE_CHECK(!keep || m2 + pos < m);
                if (keep) store_meta(other + m2 + pos, cm);
                m2 += cnt;
            }
            __syncthreads();
            CandMeta* t = cur;
            cur = other;
            other = t;
            m = m2;
            msa = build_msa<!kGated>(P, L, dims, cur, m, la, margin, &sMin, &sMax);
        }

        // ---- phase D: anchor correction (src/msa.cpp:304-453, include/corrector.hpp:1423-1472) ----
        // the forest path corrects even when refinement removed every
        // candidate (correctAnchorClf sets isCorrected unconditionally)
        const bool forestA = !kGated && P.correctionType == 1;
        if (m > 0 || forestA) {
            __shared__ float sMinSupport, sAvgSupport;
            __shared__ int sMinCov, sMaxCov, sHQ, sByCons;
            ColumnStats cs0{};
            if (wave == 0) cs0 = column_stats_wave(L, msa.ab, msa.ab + la);
            if (tid == 0) {
                const ColumnStats cs = cs0;
                const float mn = cs.minSupport, avg = cs.avgSupport;
                const int mc = cs.minCov;
                const bool allGood = fgeq(avg, P.avgSupportThreshold) && fgeq(mn, P.minSupportThreshold) &&
                                     fgeq(float(mc), P.minCoverageThreshold);
                int hq = 0;
                if (allGood) {
                    int smallest = 100;
                    for (int percent = P.errPercent; percent >= 0; percent--) {
                        const float factor = float(percent) / 100.0f;
                        if (fgeq(avg, 1.0f - factor) && fgeq(mn, fmaf(-3.0f, factor, 1.0f))) smallest = percent;
                    }
                    hq = fgeq(float(mc), P.minCoverageThreshold) && fleq(float(smallest), P.errPercentHalf);
                }
                sMinSupport = mn;
                sAvgSupport = avg;
                sMinCov = mc;
                sMaxCov = cs.maxCov;
                sHQ = hq;
                sByCons = allGood;
            }
            __syncthreads();
            const bool byCons = sByCons != 0;
            const bool hq = sHQ != 0;
            ColumnStats acs;
            acs.avgSupport = sAvgSupport;
            acs.minSupport = sMinSupport;
            acs.minCov = sMinCov;
            acs.maxCov = sMaxCov;
            const bool gated = kGated;
            if (kGated && !hq) {
                // decision columns in position order
                int nd = 0;
                for (int base = 0; base < la; base += blockDim.x) {
                    const int i = base + tid;
                    const bool d = i < la && u32(L.cons[msa.ab + i]) != base_at(L.aw, i);
                    if (i < la) L.gacc[i] = 0;
                    int cnt;
                    const int pos = block_exclusive_count(d, L.scratch, &cnt);
                    CARE_CHECK(!d || nd + pos < la);
                    if (d) L.dec[nd + pos] = u16(i);
                    nd += cnt;
                }
                __syncthreads();
                bf16raw* xt = reinterpret_cast<bf16raw*>(L.gscr);                  // [32][32]
                float* zt = reinterpret_cast<float*>(L.gscr + 32 * kMlpIn * 2);   // [32]
                char* tileScr = L.gscr + 32 * kMlpIn * 2 + 32 * 4;                // 2 x kMlpTileScratch
                float* ft = reinterpret_cast<float*>(tileScr + 2 * kMlpTileScratch);  // [32][33] features
                // the tiles overlay the candidate staging area, which ends at qw
                CARE_CHECK(reinterpret_cast<char*>(ft + 32 * (kMlpIn + 1)) <= reinterpret_cast<char*>(L.qw));
                for (int g0 = 0; g0 < nd; g0 += 32) {
                    const int ng = min(32, nd - g0);
                    // thread and wave indices made opaque: the feature-tile
                    // and MLP addresses are computed here rather than hoisted
                    // out of the anchor loop (which would keep them in
                    // registers through every phase and cost a wave per SIMD)
                    int gt = tid, gw = wave;
                    asm volatile("" : "+v"(gt), "+v"(gw));
                    if (gt < 32) {
                        float* f = ft + gt * (kMlpIn + 1);
                        if (gt < ng) {
                            const int i = int(L.dec[g0 + gt]);
                            const float qwi = P.useQuality ? L.qw[(unsigned char)L.anchorQ[i]] : 1.0f;
                            gate_features(L, msa.cols, msa.ab, la, i, nd, m, acs, P.estimatedCoverage, qwi,
                                          [&](int ii) { return base_at(L.aw, ii); }, f);
                        } else {
                            for (int k = 0; k < kMlpIn; k++) f[k] = 0.0f;
                        }
                        for (int k = 0; k < kMlpIn; k++) xt[gt * kMlpIn + k] = f2bf(f[k]);
                        if (P.gateFeat && !P.gateRecordCands && gt < ng) {
                            const int slot = atomicAdd(P.gateCount, 1);
                            if (slot < P.gateCap) {
                                const int i = int(L.dec[g0 + gt]);
                                for (int k = 0; k < kMlpIn; k++) P.gateFeat[u64(slot) * kMlpIn + k] = f[k];
                                P.gateIds[slot] = u32(id);
                                P.gatePos[slot] = u16(i);
                                P.gateOrig[slot] = u8(base_at(L.aw, i));
                                P.gateCons[slot] = L.cons[msa.ab + i];
                            }
                        }
                    }
                    __syncthreads();
                    if (P.gateParams && wave < 2 && 16 * wave < ng) {
                        MlpWeights W;
                        W.w1 = P.gateW1;
                        W.w2 = P.gateW2;
                        W.w3 = P.gateW3;
                        W.b1 = P.gateParams + kMlpB1;
                        W.b2 = P.gateParams + kMlpB2;
                        W.b3 = P.gateParams + kMlpB3;
                        W.w4 = P.gateParams + kMlpW4;
                        W.b4 = P.gateParams[kMlpB4];
                        mlp_tile_forward(W, xt + 16 * gw * kMlpIn, tileScr + gw * kMlpTileScratch, zt + 16 * gw);
                    }
                    __syncthreads();
                    if (tid < ng) L.gacc[L.dec[g0 + tid]] = (P.gateParams && zt[tid] >= P.gateLogit) ? 1 : 0;
                    __syncthreads();
                }
            }
            for (int i = tid; i < la; i += blockDim.x) {
                const int col = msa.ab + i;
                const u32 orig = base_at(L.aw, i);
                u32 b;
                if (forestA) {
                    b = L.cons[col];
                    if (!hq && b != orig) {
                        float f[14];
                        column_features(L, msa.cols, col, orig, acs, P.estimatedCoverage, f);
                        if (!forest_decide(P.forestA, f)) b = orig;
                    }
                } else if (gated) {
                    b = L.cons[col];
                    if (!hq && b != orig && !L.gacc[i]) b = orig;
                } else if (byCons) {
                    b = L.cons[col];
                } else {
                    b = (L.support[col] > 0.90f && L.origCov[col] <= 2) ? u32(L.cons[col]) : orig;
                }
                outSeq[i] = base_char(b);
            }
            status = u8(1 | (hq ? 2 : 0));

            // ---- phase E: candidate correction (src/msa.cpp:457-522,
            // include/corrector.hpp:1539-1623) ----
            if (hq && P.correctCandidates) {
                const int ae = msa.ab + la;
                const bool forestC = P.correctionTypeCands == 1;
                // learned gate on candidate columns (gated mode only): the
                // forest's candidate range and decision rule, the MLP deciding
                const bool gateC = kGateCands && P.correctionTypeCands == 3;
                for (int c = wave; c < m; c += nwaves) {
                    const CandMeta cm = cur[c];
                    const int qb = msa.ab + cm.shift;
                    const int qe = qb + cm.len;
                    const bool inRange = (forestC || gateC)
                                             ? (qb >= msa.ab - P.newColumns && qe <= ae + P.newColumns)
                                             : (msa.ab - P.newColumns <= qb && qb <= msa.ab + P.newColumns &&
                                                qe <= ae + P.newColumns);
                    if (!inRange) continue;
                    int slot = 0;
                    if (lane == 0) slot = atomicAdd(P.candOutCount, 1);
                    slot = __shfl(slot, 0, kWave);
                    if (slot >= P.candOutCap) continue;
                    CARE_CHECK(cm.id < rs.n && qb >= 0 && qe <= msa.cols);
                    if (lane == 0) P.candOutIds[slot] = cm.id;
                    ColumnStats ccs{};
                    if (forestC || gateC) {
                        ccs = column_stats_wave(L, qb, qe);
                        ccs.avgSupport = __uint_as_float(
                            u32(__builtin_amdgcn_readlane(int(__float_as_uint(ccs.avgSupport)), 0)));
                    }
                    // jaccard of the anchor and candidate column ranges
                    const float jac = float(min(ae, qe) - max(msa.ab, qb)) / float(max(ae, qe) - min(msa.ab, qb));
                    char* dst = P.candOutSeq + u64(slot) * P.outPitch;
                    const bool rc = cm.flag == 2;
                    const u32* enc = rs.seqs + u64(cm.id) * rs.pitch;
                    u8* cacc = nullptr;
                    if (gateC) {
                        char* ws = L.gscr + size_t(wave) * cand_gate_scratch(dims.maxLen);
                        CARE_CHECK(ws + cand_gate_scratch(dims.maxLen) <= reinterpret_cast<char*>(L.qw));
                        cacc = reinterpret_cast<u8*>(ws + kCandGateTile);
                        gate_candidate(P, L, rs, msa.cols, qb, cm, m, ccs, ws, cacc);
                    }
                    for (int i = lane; i < cm.len; i += kWave) {
                        // position i of the candidate in MSA orientation, written
                        // to its forward-strand position
                        const int col = qb + i;
                        u32 b = L.cons[col];
                        if (forestC) {
                            const u32 orig = oriented_base(enc, cm.len, i, rc);
                            if (b != orig) {
                                float f[14];
                                column_features(L, msa.cols, col, orig, ccs, P.estimatedCoverage, f);
                                f[13] = jac;
                                if (!forest_decide(P.forestC, f)) b = orig;
                            }
                        } else if (gateC) {
                            const u32 orig = oriented_base(enc, cm.len, i, rc);
                            if (b != orig && !cacc[i]) b = orig;
                        }
                        const int out = rc ? cm.len - 1 - i : i;
                        dst[out] = base_char(rc ? 3u - b : b);
                    }
                }
            }
        }
        if (tid == 0) P.outStatus[a] = status;
        __syncthreads();
    }
}

// LDS shape of k_correct; MSA chunks hold up to 256 candidates and about
// 16 KiB of staged rows
// chunk: candidates staged per MSA pass (0 = about 16 KB of staged rows)
static LdsDims make_dims(int maxLen, int useQuality, int qpitch, int threads, int gated, int chunk) {
    LdsDims d;
    d.pitch = (maxLen + 15) / 16;
    d.qpitch = useQuality ? qpitch : 0;
    d.hw = (maxLen + 31) / 32;
    d.maxLen = maxLen;
    d.cols = 3 * maxLen + 2;
    d.waves = threads / kWave;
    // LDS per staged candidate: record, 2-bit and quality rows, bit planes
    const int rowBytes = int(sizeof(CandMeta)) + 4 * (d.pitch + d.qpitch) + 2 * int(sizeof(uint2)) * d.pw();
    d.chunk = chunk > 0 ? chunk : std::max(16, std::min(256, 16384 / rowBytes));
    d.gated = gated;
    if (gated) {
        // the staging area doubles as the gates' feature tiles and MLP scratch
        // (bit 0: anchor gate, two waves; bit 1: candidate gate, every wave)
        int need = 0;
        if (gated & 1) need = 32 * kMlpIn * 2 + 32 * 4 + 2 * kMlpTileScratch + 32 * (kMlpIn + 1) * 4;
        if (gated & 2) need = std::max(need, d.waves * cand_gate_scratch(maxLen));
        while (d.chunk * rowBytes < need) d.chunk++;
    }
    return d;
}

size_t correct_lds_bytes(int maxLen, int useQuality, int qpitch, int threads, int gated, int chunk) {
    return make_dims(maxLen, useQuality, qpitch, threads, gated, chunk).bytes();
}

int correct_msa_chunk(int maxLen, int useQuality, int qpitch, int threads, int gated, int chunk) {
    return make_dims(maxLen, useQuality, qpitch, threads, gated, chunk).chunk;
}

int correct_resident_blocks(int threads, size_t ldsBytes, int gated) {
    int nb = 0;
    const hipError_t e =
        (gated & 2)   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_correct<true, true>, threads, ldsBytes)
        : (gated & 4) ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_correct<true, false, true>, threads,
                                                                     ldsBytes)
        : (gated & 1) ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_correct<true, false>, threads, ldsBytes)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_correct<false, false>, threads, ldsBytes);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

int correct_gate_mode(int correctionType, int correctCandidates, int correctionTypeCands) {
    return (correctionType == 3 ? 1 : 0) | (correctionType == 3 && correctCandidates && correctionTypeCands == 3 ? 2 : 0);
}

void launch_correct(const CorrectParams& p, int blocks, int threads, int chunk, hipStream_t stream) {
    const int gm = correct_gate_mode(p.correctionType, p.correctCandidates, p.correctionTypeCands);
    const LdsDims d = make_dims(p.rs.maxLen, p.useQuality, p.rs.qpitch, threads, gm, chunk);
    // the candidate gate has its own instantiation: its registers would lower
    // the occupancy of the anchor-only gated mode
    if (gm & 2)
        hipLaunchKernelGGL((k_correct<true, true>), dim3(blocks), dim3(threads), d.bytes(), stream, p, d);
    else if ((gm & 1) && p.gatedRefine)
        hipLaunchKernelGGL((k_correct<true, false, true>), dim3(blocks), dim3(threads), d.bytes(), stream, p, d);
    else if (gm & 1)
        hipLaunchKernelGGL((k_correct<true, false>), dim3(blocks), dim3(threads), d.bytes(), stream, p, d);
    else
        hipLaunchKernelGGL((k_correct<false, false>), dim3(blocks), dim3(threads), d.bytes(), stream, p, d);
}

}  // namespace care_engine
