// CPU emulation of the HIP runtime/device subset used by care_amd/csrc, so
// the engine's kernel sources can run in this container's tests without a
// GPU.  Test infrastructure only: it is compiled into a separate library
// (tests/emu/build_emu.sh) that the CPU tests load; the product is never
// built against it.
//
// Model: a kernel launch runs its blocks one after another; the threads of
// a block are fibers (ucontext) on the calling OS thread.  __syncthreads()
// is a barrier over the block, wave (64-lane) operations exchange values
// through a per-wave buffer between two wave barriers.  A scheduling pass
// in which no fiber can proceed is reported as a deadlock.  __shared__
// variables become function statics (one block runs at a time) and dynamic
// LDS is one buffer per launch.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <csetjmp>
#include <ucontext.h>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))
#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(emu::dyn_lds());

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint2 {
    unsigned x, y;
};
inline uint2 make_uint2(unsigned a, unsigned b) { return uint2{a, b}; }
struct alignas(16) uint4 {
    unsigned x, y, z, w;
};

typedef int hipError_t;
typedef void* hipStream_t;
typedef struct emu_event* hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
enum { hipStreamNonBlocking = 1, hipHostMallocDefault = 0, hipEventDisableTiming = 2 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };

namespace emu {

constexpr int kWaveSize = 64;

struct ThreadCtx {
    dim3 tid, bid, bdim, gdim;
};
// All emulated GPU threads are fibers on one OS thread; ctx is the running
// fiber's context.
inline ThreadCtx ctx;

struct Fiber {
    ucontext_t uc;
    jmp_buf jb;
    bool started = false;
    std::vector<char> stack;
    ThreadCtx c;
    bool done = false;
    int waitKind = 0;  // 0 running, 1 block barrier, 2 wave barrier
    unsigned long waitGen = 0;
};

// Lane values of a wave operation go to buffer (generation & 1): a fiber
// can only write the same buffer again two wave barriers later, after every
// lane has read it, so one barrier per operation suffices.
struct Wave {
    unsigned long long vals[2][kWaveSize];
    unsigned short mf[2][kWaveSize][16];  // matrix operands of the lanes
    int arrived = 0;
    int size = 0;
    unsigned long gen = 0;
};

struct Sched {
    std::vector<Fiber> fibers;
    std::vector<Wave> waves;
    ucontext_t main;
    jmp_buf mainJb;
    int cur = -1;
    int blockArrived = 0;
    unsigned long blockGen = 0;
    int blockSite = 0;
    void (*body)(void*) = nullptr;
    void* arg = nullptr;
};

inline Sched*& sched() {
    static Sched* s = nullptr;
    return s;
}
inline std::vector<char>& lds() {
    static std::vector<char> b;
    return b;
}
inline char* dyn_lds() { return lds().data(); }
// dynamic LDS bytes of the running launch
inline size_t& dyn_lds_bytes() {
    static size_t n = 0;
    return n;
}
// EMU_CUS=<n>: compute units reported by the device attribute query
// (default 2; 256 gives the launch grids of an MI355X)
inline int num_cus() {
    static const int v = [] {
        const char* e = std::getenv("EMU_CUS");
        const int n = e ? std::atoi(e) : 2;
        return n > 0 ? n : 2;
    }();
    return v;
}
// kernel index checks (CARE_CHECK in the kernel sources, built with
// CARE_EMU_CHECKS): a failed check aborts with its location
inline void check(bool ok, const char* what, const char* file, int line);
inline int lane() { return int(ctx.tid.x % kWaveSize); }
inline Wave& my_wave() { return sched()->waves[ctx.tid.x / kWaveSize]; }

// Fibers enter through makecontext/swapcontext once; later switches use
// _setjmp/_longjmp, which do not save the signal mask (no system call).
inline void yield_to_scheduler() {
    Sched* s = sched();
    Fiber& f = s->fibers[s->cur];
    if (_setjmp(f.jb) == 0) _longjmp(s->mainJb, 1);
}

// site: source line of the __syncthreads() call in checked builds.  The
// threads of a block must meet at one call site: a barrier that pairs
// different sites (one reached by a subset of the waves) is reported.
inline void block_barrier(int site = 0) {
    Sched* s = sched();
    Fiber& f = s->fibers[s->cur];
    const unsigned long g = s->blockGen;
    if (s->blockArrived == 0) {
        s->blockSite = site;
    } else if (site != s->blockSite) {
        std::fprintf(stderr, "emu: threads of block %u meet at different barriers (lines %d and %d, thread %u)\n",
                     ctx.bid.x, s->blockSite, site, ctx.tid.x);
        std::abort();
    }
    if (++s->blockArrived == int(s->fibers.size())) {
        s->blockArrived = 0;
        s->blockGen++;
        return;
    }
    f.waitKind = 1;
    f.waitGen = g;
    yield_to_scheduler();
}

inline void wave_barrier() {
    Sched* s = sched();
    Fiber& f = s->fibers[s->cur];
    Wave& w = my_wave();
    const unsigned long g = w.gen;
    if (++w.arrived == w.size) {
        w.arrived = 0;
        w.gen++;
        return;
    }
    f.waitKind = 2;
    f.waitGen = g;
    yield_to_scheduler();
}

template <class T>
inline unsigned long long to_bits(T v) {
    unsigned long long b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T from_bits(unsigned long long b) {
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}
template <class T>
inline T exchange(T v, int src) {
    Wave& w = my_wave();
    const unsigned long g = w.gen;
    w.vals[g & 1][lane()] = to_bits(v);
    wave_barrier();
    return from_bits<T>(w.vals[g & 1][src]);
}

// 16x16 output tile of a BF16 matrix instruction with K = 4 * per: lane l
// holds A[l%16][per*(l/16) ..], B[per*(l/16) ..][l%16] and
// D[4(l/16) .. +3][l%16]; products summed in FP32 in k order.
typedef float emu_f4 __attribute__((vector_size(16)));
inline emu_f4 mfma_bf16(const unsigned short* a, const unsigned short* b, int per, emu_f4 c) {
    Wave& w = my_wave();
    const unsigned long g = w.gen;
    const int l = lane();
    for (int j = 0; j < per; j++) {
        w.mf[g & 1][l][j] = a[j];
        w.mf[g & 1][l][8 + j] = b[j];
    }
    wave_barrier();
    auto f = [](unsigned short v) {
        const unsigned u = unsigned(v) << 16;
        float x;
        std::memcpy(&x, &u, 4);
        return x;
    };
    emu_f4 d;
    for (int i = 0; i < 4; i++) {
        const int row = 4 * (l / 16) + i, col = l % 16;
        float acc = 0.f;
        for (int k = 0; k < 4 * per; k++) {
            const float av = f(w.mf[g & 1][row + 16 * (k / per)][k % per]);
            const float bv = f(w.mf[g & 1][col + 16 * (k / per)][8 + k % per]);
            acc += av * bv;
        }
        d[i] = c[i] + acc;
    }
    return d;
}

inline void fiber_entry() {
    Sched* s = sched();
    s->body(s->arg);
    s->fibers[s->cur].done = true;
    _longjmp(s->mainJb, 1);
}

// Kernels of the engine accept any block size that is a multiple of 64;
// EMU_BLOCK (default 256) caps the emulated block size.
inline unsigned block_cap() {
    const char* e = std::getenv("EMU_BLOCK");
    const int v = e ? std::atoi(e) : 256;
    return unsigned(v >= 64 ? v : 64);
}

// Runs one block to completion; aborts with a report on deadlock.
inline void run_block(Sched& s, const char* name) {
    const int n = int(s.fibers.size());
    for (int i = 0; i < n; i++) {
        Fiber& f = s.fibers[i];
        if (f.stack.size() < 128 * 1024) f.stack.resize(128 * 1024);  // kept across launches
        getcontext(&f.uc);
        f.uc.uc_stack.ss_sp = f.stack.data();
        f.uc.uc_stack.ss_size = f.stack.size();
        f.uc.uc_link = nullptr;
        makecontext(&f.uc, fiber_entry, 0);
        f.started = false;
    }
    // EMU_SCHED=reverse|random: order in which runnable fibers are resumed
    // in every pass (default: thread order), to vary the interleaving of
    // waves between barriers.
    static const int schedMode = [] {
        const char* e = std::getenv("EMU_SCHED");
        if (!e) return 0;
        return std::strcmp(e, "reverse") == 0 ? 1 : std::strcmp(e, "random") == 0 ? 2 : 0;
    }();
    static unsigned long long rng = 0x9E3779B97F4A7C15ull;
    std::vector<int> order(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) order[size_t(i)] = schedMode == 1 ? n - 1 - i : i;
    for (;;) {
        bool allDone = true, progress = false;
        if (schedMode == 2) {
            for (int i = n - 1; i > 0; i--) {
                rng ^= rng << 13;
                rng ^= rng >> 7;
                rng ^= rng << 17;
                std::swap(order[size_t(i)], order[size_t(rng % unsigned(i + 1))]);
            }
        }
        for (int k = 0; k < n; k++) {
            const int i = order[size_t(k)];
            Fiber& f = s.fibers[i];
            if (f.done) continue;
            allDone = false;
            if (f.waitKind == 1 && f.waitGen == s.blockGen) continue;
            if (f.waitKind == 2 && f.waitGen == s.waves[f.c.tid.x / kWaveSize].gen) continue;
            f.waitKind = 0;
            s.cur = i;
            ctx = f.c;
            if (_setjmp(s.mainJb) == 0) {
                if (!f.started) {
                    f.started = true;
                    swapcontext(&s.main, &f.uc);
                } else {
                    _longjmp(f.jb, 1);
                }
            }
            progress = true;
        }
        if (allDone) break;
        if (!progress) {
            std::fprintf(stderr, "emu: deadlock in kernel %s block %u:", name, s.fibers[0].c.bid.x);
            for (int i = 0; i < n; i++)
                std::fprintf(stderr, " t%d:%s", i, s.fibers[i].done ? "done" : s.fibers[i].waitKind == 1 ? "block" : "wave");
            std::fprintf(stderr, "\n");
            std::abort();
        }
    }
}

inline Sched& sched_storage() {
    static Sched s;
    return s;
}

// EMU_POISON=<byte>: fill device allocations and, before every block, the
// dynamic LDS with that byte instead of zeros, so reads of memory that no
// kernel or copy wrote show up as changed results.
inline int poison_byte() {
    static const int v = [] {
        const char* e = std::getenv("EMU_POISON");
        return e ? int(std::strtol(e, nullptr, 0) & 0xFF) : -1;
    }();
    return v;
}

// Stream capture: while a capture is open, launches and copies are recorded
// (arguments by value) instead of run; launching the graph runs them in
// order.
inline std::vector<std::function<void()>>*& capture_list() {
    static std::vector<std::function<void()>>* c = nullptr;
    return c;
}

template <class K, class... Args>
void launch_run(const char* name, K kernel, dim3 grid, dim3 block, size_t ldsBytes, Args... args);

template <class K, class... Args>
void launch_named(const char* name, K kernel, dim3 grid, dim3 block, size_t ldsBytes, Args... args) {
    if (capture_list()) {
        capture_list()->push_back([=]() { launch_run(name, kernel, grid, block, ldsBytes, args...); });
        return;
    }
    launch_run(name, kernel, grid, block, ldsBytes, args...);
}

template <class K, class... Args>
void launch_run(const char* name, K kernel, dim3 grid, dim3 block, size_t ldsBytes, Args... args) {
    if (block.x > block_cap()) block.x = block_cap();
    const unsigned nthreads = block.x;
    lds().assign(ldsBytes ? ldsBytes : 1, 0);
    dyn_lds_bytes() = ldsBytes;
    auto call = [&]() { kernel(args...); };
    Sched& s = sched_storage();
    s.body = [](void* p) { (*static_cast<decltype(call)*>(p))(); };
    s.arg = &call;
    sched() = &s;
    s.fibers.resize(nthreads);
    for (unsigned b = 0; b < grid.x; b++) {
        if (poison_byte() >= 0) std::memset(lds().data(), poison_byte(), lds().size());
        for (auto& f : s.fibers) {
            f.done = false;
            f.waitKind = 0;
            f.waitGen = 0;
        }
        s.blockGen = 0;
        const unsigned nw = (nthreads + kWaveSize - 1) / kWaveSize;
        s.waves.assign(nw, Wave{});
        for (unsigned w = 0; w < nw; w++) s.waves[w].size = int(std::min<unsigned>(kWaveSize, nthreads - w * kWaveSize));
        s.blockArrived = 0;
        for (unsigned t = 0; t < nthreads; t++) {
            s.fibers[t].c.tid = dim3(t);
            s.fibers[t].c.bid = dim3(b);
            s.fibers[t].c.bdim = block;
            s.fibers[t].c.gdim = grid;
        }
        run_block(s, name);
    }
    sched() = nullptr;
}

}  // namespace emu

inline void emu::check(bool ok, const char* what, const char* file, int line) {
    if (ok) return;
    std::fprintf(stderr, "emu: check failed: %s at %s:%d (block %u thread %u)\n", what, file, line, ctx.bid.x,
                 ctx.tid.x);
    std::abort();
}

#define threadIdx (emu::ctx.tid)
#define blockIdx (emu::ctx.bid)
#define blockDim (emu::ctx.bdim)
#define gridDim (emu::ctx.gdim)

#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
    emu::launch_named(#kernel, kernel, grid, block, lds, __VA_ARGS__)

#ifdef CARE_EMU_CHECKS
#define __syncthreads() emu::block_barrier(__LINE__)
#else
inline void __syncthreads() { emu::block_barrier(); }
#endif
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_wave_barrier() emu::wave_barrier()

inline unsigned __builtin_amdgcn_alignbit(unsigned a, unsigned b, unsigned c) {
    return unsigned((((unsigned long long)a << 32) | b) >> (c & 31));
}
inline unsigned __lane_id() { return unsigned(emu::lane()); }
inline unsigned long long __ballot(int pred) {
    emu::Wave& w = emu::my_wave();
    const unsigned long g = w.gen;
    w.vals[g & 1][emu::lane()] = pred ? 1ull : 0ull;
    emu::wave_barrier();
    unsigned long long m = 0;
    for (int i = 0; i < w.size; i++)
        if (w.vals[g & 1][i]) m |= 1ull << i;
    return m;
}
inline int __any(int pred) { return __ballot(pred) != 0ull; }
#define __bf16 unsigned short
typedef unsigned short emu_us8 __attribute__((vector_size(16)));
typedef short emu_s4 __attribute__((vector_size(8)));
inline emu::emu_f4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(emu_us8 a, emu_us8 b, emu::emu_f4 c, int, int, int) {
    unsigned short av[8], bv[8];
    for (int j = 0; j < 8; j++) {
        av[j] = a[j];
        bv[j] = b[j];
    }
    return emu::mfma_bf16(av, bv, 8, c);
}
inline emu::emu_f4 __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(emu_s4 a, emu_s4 b, emu::emu_f4 c, int, int, int) {
    unsigned short av[4], bv[4];
    for (int j = 0; j < 4; j++) {
        av[j] = (unsigned short)a[j];
        bv[j] = (unsigned short)b[j];
    }
    return emu::mfma_bf16(av, bv, 4, c);
}
inline int __all(int pred) {
    emu::Wave& w = emu::my_wave();
    const unsigned long long full = w.size >= 64 ? ~0ull : ((1ull << w.size) - 1);
    return __ballot(pred) == full;
}
template <class T>
inline T __shfl(T v, int src, int width = 64) {
    (void)width;
    return emu::exchange(v, src & 63);
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    (void)width;
    return emu::exchange(v, (emu::lane() ^ mask) & 63);
}
// DPP moves used by the kernels (source lane within the 16-lane row):
// quad_perm 0x00-0xFF, row_mirror 0x140, row_half_mirror 0x141
inline int __builtin_amdgcn_mov_dpp(int v, int ctrl, int, int, bool) {
    const int l = emu::lane(), row = l & ~15, i = l & 15;
    int src;
    if (ctrl >= 0 && ctrl <= 0xFF) src = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
    else if (ctrl == 0x140) src = row | (15 - i);
    else if (ctrl == 0x141) src = row | (i & 8) | (7 - (i & 7));
    else {
        std::fprintf(stderr, "emulated mov_dpp: control 0x%x not emulated\n", ctrl);
        std::abort();
    }
    return emu::exchange(v, src);
}
inline int __builtin_amdgcn_readlane(int v, int lane) { return emu::exchange(v, lane & 63); }
// every lane of an emulated wave reaches the call (no exec mask), so the
// first active lane is lane 0
inline int __builtin_amdgcn_readfirstlane(int v) { return emu::exchange(v, 0); }
template <class T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
    (void)width;
    const int src = emu::lane() - int(delta);
    const T r = emu::exchange(v, src < 0 ? emu::lane() : src);
    return r;
}

inline float __uint_as_float(unsigned x) {
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}
inline unsigned __float_as_uint(float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    return u;
}
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz(unsigned(x)); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
inline unsigned __brev(unsigned x) {
    unsigned r = 0;
    for (int i = 0; i < 32; i++) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
using std::max;
using std::min;

template <class T>
inline T atomicAdd(T* p, T v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicMin(int* p, int v) {
    int cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
inline unsigned atomicMin(unsigned* p, unsigned v) {
    unsigned cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
inline int atomicMax(int* p, int v) {
    int cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
inline int atomicExch(int* p, int v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicCAS(unsigned long long* p, unsigned long long expected, unsigned long long desired) {
    __atomic_compare_exchange_n(p, &expected, desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return expected;
}

// runtime API
struct emu_event {
    double t;
};
inline hipError_t hipGetDeviceCount(int* c) {
    *c = 1;
    return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = reinterpret_cast<hipStream_t>(1);
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
// launches and copies complete when issued, so cross-stream waits are no-ops
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new emu_event{0};
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    e->t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = float(b->t - a->t);
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    *v = emu::num_cus();
    return hipSuccess;
}
// blocks resident per compute unit: bounded by 64 KB of LDS and 8 blocks
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* nb, K, int, size_t lds) {
    const size_t byLds = (64 * 1024) / (lds ? lds : 1);
    *nb = int(byLds < 8 ? (byLds ? byLds : 1) : 8);
    return hipSuccess;
}
// device allocations are 256-byte aligned, as on the device
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
    const size_t n = (bytes + 255) / 256 * 256 + 256 * (bytes == 0);
    void* v = std::aligned_alloc(256, n);
    if (!v) return hipErrorInvalidValue;
    std::memset(v, emu::poison_byte() >= 0 ? emu::poison_byte() : 0, n);
    *p = static_cast<T*>(v);
    return hipSuccess;
}
inline hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
// page-locked allocations are remembered so that pointer queries can tell
// them from pageable memory
namespace emu {
inline std::mutex& host_alloc_mutex() {
    static std::mutex m;
    return m;
}
inline std::map<const char*, size_t>& host_allocs() {
    static std::map<const char*, size_t> m;
    return m;
}
}  // namespace emu
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
    *p = std::calloc(1, bytes ? bytes : 1);
    if (!*p) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> g(emu::host_alloc_mutex());
    emu::host_allocs()[static_cast<const char*>(*p)] = bytes ? bytes : 1;
    return hipSuccess;
}
inline hipError_t hipHostFree(void* p) {
    {
        std::lock_guard<std::mutex> g(emu::host_alloc_mutex());
        emu::host_allocs().erase(static_cast<const char*>(p));
    }
    std::free(p);
    return hipSuccess;
}
enum hipMemoryType { hipMemoryTypeUnregistered = 0, hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2 };
struct hipPointerAttribute_t {
    hipMemoryType type;
    int device;
    void* devicePointer;
    void* hostPointer;
    int isManaged;
    unsigned allocationFlags;
};
// page-locked memory from hipHostMalloc is reported as host memory, anything
// else as an error (as for pageable memory on the device runtime)
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t* a, const void* p) {
    std::lock_guard<std::mutex> g(emu::host_alloc_mutex());
    const char* c = static_cast<const char*>(p);
    auto& m = emu::host_allocs();
    auto it = m.upper_bound(c);
    if (it == m.begin()) return hipErrorInvalidValue;
    --it;
    if (c >= it->first + it->second) return hipErrorInvalidValue;
    *a = hipPointerAttribute_t{hipMemoryTypeHost, 0, const_cast<char*>(c), const_cast<char*>(c), 0, 0};
    return hipSuccess;
}
inline hipError_t hipMemcpy2DAsync(void* d, size_t dpitch, const void* s, size_t spitch, size_t width,
                                   size_t height, hipMemcpyKind, hipStream_t = nullptr) {
    auto run = [=]() {
        for (size_t r = 0; r < height; r++)
            std::memcpy(static_cast<char*>(d) + r * dpitch, static_cast<const char*>(s) + r * spitch, width);
    };
    if (emu::capture_list()) emu::capture_list()->push_back(run);
    else run();
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    if (emu::capture_list()) emu::capture_list()->push_back([=]() { std::memcpy(d, s, n); });
    else std::memcpy(d, s, n);
    return hipSuccess;
}
// graphs: the recorded work of one capture
struct emu_graph {
    std::vector<std::function<void()>> ops;
};
typedef emu_graph* hipGraph_t;
typedef emu_graph* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (emu::capture_list()) return hipErrorInvalidValue;
    emu::capture_list() = new std::vector<std::function<void()>>();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    if (!emu::capture_list()) return hipErrorInvalidValue;
    *g = new emu_graph{std::move(*emu::capture_list())};
    delete emu::capture_list();
    emu::capture_list() = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new emu_graph{g->ops};
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    if (emu::capture_list()) return hipErrorInvalidValue;
    for (auto& op : e->ops) op();
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
// the address of a __device__ variable (host memory in the emulation)
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
// a device with 8 GiB free
inline hipError_t hipMemGetInfo(size_t* freeB, size_t* totalB) {
    *freeB = size_t(8) << 30;
    *totalB = size_t(16) << 30;
    return hipSuccess;
}
inline const char* hipGetErrorString(hipError_t) { return "emulated HIP error"; }
